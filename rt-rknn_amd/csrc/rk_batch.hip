// rk_batch.hip -- batched form: many queries with uneven candidate counts and
// mixed k in one hipGraph.
//
// The graph is a fixed launch sequence over device-side work lists:
//   prep    zero the counters; candidate records for every (query, candidate),
//           binary64 -> fp32
//   root    one wave per query tests the <= 64 top tree nodes and lists
//           the survivors
//   level p = top .. 1
//           items (query, node at level p); the node's 64 children on lanes;
//           children not certainly covered by k candidates are listed below
//   leaf    items (query, tile); per-user count as in the plain form;
//           survivors go straight into per-query result segments
//   sort    per-query block radix sort of each result segment
// The fetch compacts the segments (offsets scan + copy) outside the graph.
// Every kernel is a grid-stride loop over a device-side item count, so the
// graph is captured once and replayed unchanged for any data; nothing waits
// on another wave, so every launch drains.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "rk_internal.hpp"
#include "rknn.h"

using namespace rk;

namespace {

constexpr uint64_t kItemBudget = 1ull << 25; // per level
constexpr uint64_t kPoolBudget = 1ull << 26; // result ids per batch (all segments)
constexpr int kWavesPerCU = 32;             // persistent grids: 8 waves per SIMD

rk_status build_graph(rk_batch* b)
{
    rk_dataset* ds = b->ds;
    rk_ctx* c = ds->ctx;
    hipStream_t st = c->stream;
    const int grid = device_grid(c->n_cu, kWavesPerCU);
    RK_HIP(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    // the root kernel lists each query's surviving top nodes; the level
    // kernels then descend level by level down to the tiles
    launch_batch_prep(b->view, b->cand.p, ds->fx.p, ds->fy.p, b->wb.p, b->total_cand, st);
    launch_batch_root(ds->view, b->view, grid, st);
    for (uint32_t p = ds->view.n_levels - 1; p >= 1; --p) launch_batch_internal(ds->view, b->view, p, grid, st);
    launch_batch_leaf(ds->view, b->view, grid, st);
    launch_batch_sort(b->view, st);
    hipError_t e = hipStreamEndCapture(st, &b->graph);
    if (e != hipSuccess) return fail_hip(c, e, "hipStreamEndCapture", __LINE__);
    RK_HIP(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return RK_OK;
}

} // namespace

extern "C" {

rk_status rk_batch_prepare(rk_dataset* ds, const rk_query_desc* qs, size_t nq, rk_batch** out)
{
    if (!ds || !out || (nq && !qs)) return RK_ERR_INVALID_ARG;
    *out = nullptr;
    rk_ctx* c = ds->ctx;
    if (nq > (1u << 24)) return fail(c, RK_ERR_INVALID_ARG, "too many queries in one batch");
    std::unique_ptr<rk_batch> b(new rk_batch());
    b->ds = ds;
    b->nq = (uint32_t)nq;
    b->hq.resize(nq);
    std::vector<uint32_t> all;
    std::vector<double2> wb(nq);
    std::vector<uint32_t> cand, cl;
    const double tf0 = now_s();
    double filt = 0.0;
    for (size_t i = 0; i < nq; ++i) {
        const rk_query_desc& d = qs[i];
        if (d.q >= ds->nf) return fail(c, RK_ERR_INVALID_ARG, "query index q must be < number of facilities");
        if (d.k == 0) return fail(c, RK_ERR_INVALID_ARG, "k must be a positive integer");
        rk_status s;
        if (d.cand == nullptr && d.nc == 0) {
            const double f0 = now_s();
            ds->index->filter(d.q, d.k, cand);
            filt += now_s() - f0;
            s = clean_candidates(ds, d.q, cand.data(), cand.size(), cl);
        } else {
            s = clean_candidates(ds, d.q, d.cand, d.nc, cl);
        }
        if (s != RK_OK) return s;
        QDev& h = b->hq[i];
        h.qx = ds->hfx[d.q];
        h.qy = ds->hfy[d.q];
        h.k = d.k;
        h.nc = (uint32_t)cl.size();
        h.off = all.size();
        h.trivial = (d.k > cl.size() || ds->nu == 0) ? 1u : 0u;
        double wx = 0, wy = 0;
        if (ds->nu) axis_bounds(ds->ubox[0], ds->ubox[1], ds->ubox[2], ds->ubox[3], h.qx, h.qy, &wx, &wy);
        wb[i] = make_double2(wx, wy);
        all.insert(all.end(), cl.begin(), cl.end());
    }
    (void)tf0;
    b->filter_s = filt;
    b->total_cand = all.size();

    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    RK_HIP(c, b->q.reserve(nq));
    RK_HIP(c, b->wb.reserve(nq));
    RK_HIP(c, b->cand.reserve(all.size()));
    RK_HIP(c, b->rec.reserve(all.size()));
    RK_HIP(c, b->orig.reserve(all.size()));
    RK_HIP(c, b->qcount.reserve(nq));
    RK_HIP(c, b->qoff.reserve(nq + 1));
    RK_HIP(c, b->ctr.reserve(1));
    RK_HIP(c, b->h_ctr.reserve(1));
    // work lists of tree levels 0 .. top
    for (uint32_t p = 0; p < ds->view.n_levels; ++p) {
        uint64_t cap = (uint64_t)nq * ds->view.nlev[p];
        cap = std::max<uint64_t>(1, std::min(cap, kItemBudget));
        uint64_t bcap = (cap + kBuckets - 1) / kBuckets;
        // a bucket also takes one wave's whole staged flush
        bcap = std::max<uint64_t>(bcap, 256);
        RK_HIP(c, b->items[p].reserve(bcap * kBuckets));
        b->view.items[p] = b->items[p].p;
        b->view.item_cap[p] = (uint32_t)bcap;
    }
    // result segment per query: the whole budget split evenly, never more than the users
    uint64_t seg_cap = std::max<uint64_t>(1, std::min<uint64_t>(ds->nu, nq ? kPoolBudget / nq : 1));
    seg_cap = std::min<uint64_t>(seg_cap, 0x7fffffffull);
    RK_HIP(c, b->seg.reserve(std::max<uint64_t>(1, nq * seg_cap)));
    if (nq) RK_HIP(c, hipMemcpyAsync(b->q.p, b->hq.data(), nq * sizeof(QDev), hipMemcpyHostToDevice, st));
    if (nq) RK_HIP(c, hipMemcpyAsync(b->wb.p, wb.data(), nq * sizeof(double2), hipMemcpyHostToDevice, st));
    if (!all.empty())
        RK_HIP(c, hipMemcpyAsync(b->cand.p, all.data(), all.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    RK_HIP(c, hipStreamSynchronize(st));

    BatchDev& v = b->view;
    v.q = b->q.p;
    v.nq = b->nq;
    v.rec = b->rec.p;
    v.orig = b->orig.p;
    v.seg = b->seg.p;
    v.seg_cap = (uint32_t)seg_cap;
    v.qcount = b->qcount.p;
    v.qoff = b->qoff.p;
    v.out = b->out.p;
    v.ctr = b->ctr.p;
    if (ds->nu && nq) {
        rk_status s = build_graph(b.get());
        if (s != RK_OK) return s;
    }
    *out = b.release();
    return RK_OK;
}

rk_status rk_batch_run(rk_batch* b, double* device_seconds)
{
    if (!b) return RK_ERR_INVALID_ARG;
    rk_ctx* c = b->ds->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    b->ran = true;
    if (!b->exec) {
        if (device_seconds) *device_seconds = 0.0;
        std::memset(b->h_ctr.p, 0, sizeof(BatchCounters));
        return RK_OK;
    }
    RK_HIP(c, hipEventRecord(c->ev0, c->stream));
    RK_HIP(c, hipGraphLaunch(b->exec, c->stream));
    RK_HIP(c, hipEventRecord(c->ev1, c->stream));
    RK_HIP(c, hipMemcpyAsync(b->h_ctr.p, b->ctr.p, sizeof(BatchCounters), hipMemcpyDeviceToHost, c->stream));
    RK_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    RK_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (device_seconds) *device_seconds = ms * 1e-3;
    if (b->h_ctr.p->overflow)
        return fail(c, RK_ERR_CAPACITY, "batch exceeds the item or result budget; split it into smaller batches");
    return RK_OK;
}

rk_status rk_batch_fetch(rk_batch* b, int32_t* out_ids, size_t cap, uint64_t* offsets)
{
    if (!b || !offsets || (cap && !out_ids)) return RK_ERR_INVALID_ARG;
    rk_dataset* ds = b->ds;
    rk_ctx* c = ds->ctx;
    if (!b->ran) return fail(c, RK_ERR_INVALID_ARG, "rk_batch_run has not been called");
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    if (b->exec && b->h_ctr.p->overflow)
        return fail(c, RK_ERR_CAPACITY, "the last run overflowed its item or result budget; split the batch");
    const uint32_t nq = b->nq;
    std::vector<uint64_t> qoff(nq + 1, 0);
    if (b->exec) {
        launch_batch_scan(b->view, c->stream);
        RK_HIP(c, hipGetLastError());
        RK_HIP(c, hipMemcpy(qoff.data(), b->qoff.p, (nq + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    }
    uint64_t total = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        offsets[i] = total;
        total += b->hq[i].trivial ? (b->hq[i].k > b->hq[i].nc ? ds->nu : 0) : (qoff[i + 1] - qoff[i]);
    }
    offsets[nq] = total;
    if (total > cap) return fail(c, RK_ERR_CAPACITY, "output buffer too small; offsets[nq] holds the size needed");
    std::vector<int32_t> dev;
    if (b->exec && qoff[nq]) {
        RK_HIP(c, b->out.reserve(qoff[nq]));
        b->view.out = b->out.p; // compaction is outside the graph, so the pointer may change
        launch_batch_compact(b->view, c->stream);
        RK_HIP(c, hipGetLastError());
        dev.resize(qoff[nq]);
        RK_HIP(c, hipMemcpyAsync(dev.data(), b->out.p, qoff[nq] * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        RK_HIP(c, hipStreamSynchronize(c->stream));
    }
    for (uint32_t i = 0; i < nq; ++i) {
        int32_t* dst = out_ids + offsets[i];
        const uint64_t n = offsets[i + 1] - offsets[i];
        if (b->hq[i].trivial) {
            if (n) std::memcpy(dst, ds->all_ids_sorted.data(), n * sizeof(int32_t));
        } else if (n) {
            std::memcpy(dst, dev.data() + qoff[i], n * sizeof(int32_t));
            if (n > 4096) std::sort(dst, dst + n); // larger than the device block sort
        }
    }
    return RK_OK;
}

rk_status rk_batch_stats(const rk_batch* b, uint64_t* items, size_t n_items_cap, size_t* n_levels, uint64_t* results,
                         uint64_t* rechecks)
{
    if (!b || !b->ran) return RK_ERR_INVALID_ARG;
    const BatchCounters& c = *b->h_ctr.p;
    const size_t nl = b->ds->view.n_levels;
    if (n_levels) *n_levels = nl;
    for (size_t l = 0; l < nl && l < n_items_cap && items; ++l) {
        uint64_t t = 0;
        for (int j = 0; j < kBuckets; ++j) t += std::min(c.items[l][j], b->view.item_cap[l]);
        items[l] = t;
    }
    if (results) *results = c.results;
    if (rechecks) {
        uint64_t t = 0;
        for (int j = 0; j < kBuckets; ++j) t += c.rechecks[j];
        *rechecks = t;
    }
    return RK_OK;
}

rk_status rk_batch_destroy(rk_batch* b)
{
    if (!b) return RK_ERR_INVALID_ARG;
    rk_ctx* c = b->ds->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete b;
    return RK_OK;
}

rk_status rk_query_batch(rk_dataset* ds, const rk_query_desc* qs, size_t nq, int32_t* out_ids, size_t cap,
                         uint64_t* offsets, rk_timing* t)
{
    if (!ds || !offsets || (nq && !qs)) return RK_ERR_INVALID_ARG;
    const double t0 = now_s();
    if (t) *t = rk_timing{};
    rk_batch* b = nullptr;
    rk_status s = rk_batch_prepare(ds, qs, nq, &b);
    if (s != RK_OK) return s;
    const double t1 = now_s();
    double dev = 0.0;
    s = rk_batch_run(b, &dev);
    if (s == RK_ERR_CAPACITY && nq > 1) {
        // over the item/result budget: split in halves (results are per query)
        rk_batch_destroy(b);
        const size_t h = nq / 2;
        std::vector<uint64_t> o1(h + 1), o2(nq - h + 1);
        rk_timing ta{}, tb{};
        s = rk_query_batch(ds, qs, h, out_ids, cap, o1.data(), &ta);
        if (s != RK_OK && s != RK_ERR_CAPACITY) return s;
        const uint64_t used = o1[h];
        const bool fit1 = s == RK_OK;
        s = rk_query_batch(ds, qs + h, nq - h, fit1 && out_ids ? out_ids + used : nullptr,
                           fit1 && cap > used ? cap - used : 0, o2.data(), &tb);
        for (size_t i = 0; i <= h; ++i) offsets[i] = o1[i];
        for (size_t i = 1; i <= nq - h; ++i) offsets[h + i] = used + o2[i];
        if (t) {
            t->filter_s = ta.filter_s + tb.filter_s;
            t->kernel_s = ta.kernel_s + tb.kernel_s;
            t->n_candidates = ta.n_candidates + tb.n_candidates;
            t->n_fp64_rechecks = ta.n_fp64_rechecks + tb.n_fp64_rechecks;
            t->total_s = now_s() - t0;
        }
        if (!fit1 || s != RK_OK) return fail(ds->ctx, RK_ERR_CAPACITY, "output buffer too small");
        return RK_OK;
    }
    if (s != RK_OK) {
        rk_batch_destroy(b);
        return s;
    }
    const double t2 = now_s();
    s = rk_batch_fetch(b, out_ids, cap, offsets);
    if (t) {
        t->filter_s = b->filter_s;
        t->upload_s = t1 - t0 - b->filter_s;
        t->kernel_s = dev;
        t->download_s = now_s() - t2;
        t->n_candidates = b->total_cand;
        uint64_t rech = 0;
        for (int j = 0; j < kBuckets && b->h_ctr.p; ++j) rech += b->h_ctr.p->rechecks[j];
        t->n_fp64_rechecks = rech;
        t->total_s = now_s() - t0;
    }
    rk_batch_destroy(b);
    return s;
}

} // extern "C"
