// rk_batch.hip -- batched form: many queries with uneven candidate counts and
// mixed k in one hipGraph.
//
// A step is two launches, one block per query.  k_batch_query: candidate
// records (binary64 -> fp32), the breadth-first descent of the bounding-
// rectangle tree through LDS frontiers, the per-tile dominance count, and
// the query's result ids collected, sorted and written to the batch's
// landing area in pinned host memory; k_batch_land: the same for results
// too large for the block's LDS.  Queries share only the landing area's
// reservation counter (one atomic per query).  When the run is done, its
// results are in host memory: the fetch is a host copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "rk_internal.hpp"
#include "rknn.h"

using namespace rk;

namespace {

constexpr uint64_t kPoolBudget = 1ull << 26; // result ids per batch (all segments)

// A positive integer from the environment, or dflt.
uint64_t env_count(const char* name, uint64_t dflt)
{
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    char* end = nullptr;
    const unsigned long long x = std::strtoull(v, &end, 10);
    return (end && *end == 0 && x > 0) ? (uint64_t)x : dflt;
}

// Result ids per batch (RK_BATCH_POOL_IDS) and device candidate slots per
// query (RK_DFILTER_SLOTS, at most the default): smaller values hold less
// memory per batch, and the queries that outgrow them run again
// (batch_collect).
uint64_t pool_budget() { return env_count("RK_BATCH_POOL_IDS", kPoolBudget); }
} // namespace
namespace rk {
uint64_t batch_pool_ids() { return pool_budget(); }
} // namespace rk
namespace {
uint64_t dfilter_slots() { return std::min<uint64_t>(env_count("RK_DFILTER_SLOTS", kDfBatchCand), kDfBatchCand); }

rk_status build_graph(rk_batch* b)
{
    rk_dataset* ds = b->ds;
    rk_ctx* c = ds->ctx;
    hipStream_t st = c->stream;
    RK_HIP(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    batch_launch(b, st);
    // (the results, counts and landing offsets reach pinned host memory from
    // the step's kernels: the graph holds no copy node)
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        return fail_hip(c, e, "kernel launch (capture)", __LINE__);
    }
    e = hipStreamEndCapture(st, &b->graph);
    if (e != hipSuccess) return fail_hip(c, e, "hipStreamEndCapture", __LINE__);
    RK_HIP(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return RK_OK;
}

// A prepared batch's graph captured again (its landing area or copy size
// changed).  Caller holds the context lock; no run in flight.
rk_status recapture(rk_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return build_graph(b);
}

// The landing area holds at least `ids` ids (host and device sides, capped
// at kMaxLandIds); its reservation counters start at zero.  Caller holds the
// context lock; the batch has no run in flight.
rk_status ensure_landing(rk_batch* b, uint64_t ids)
{
    rk_ctx* c = b->ds->ctx;
    ids = std::min<uint64_t>(std::max<uint64_t>(ids, 1), kMaxLandIds);
    if (!b->land_ctr.p) {
        RK_HIP(c, b->land_ctr.reserve(2));
        RK_HIP(c, hipMemsetAsync(b->land_ctr.p, 0, 2 * sizeof(unsigned long long), c->stream));
        RK_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (!b->land_host || b->land_host->n < ids) {
        // a new host side (an old one may still be referenced by served
        // results): from the batch's source of areas when it has one
        const uint64_t want = std::min<uint64_t>(
            b->land_host ? std::max<uint64_t>(ids, b->land_host->n + b->land_host->n / 2) : ids, kMaxLandIds);
        if (b->area_source) {
            b->land_host = b->area_source(want);
            if (!b->land_host) return fail(c, RK_ERR_OUT_OF_MEMORY, "pinned landing area");
        } else {
            auto a = std::make_shared<PinnedIds>();
            RK_HIP(c, a->alloc(want));
            b->land_host = std::move(a);
        }
    }
    RK_HIP(c, b->dland.reserve(ids));
    b->land_cap = std::min<uint64_t>(std::min<uint64_t>(b->land_host->n, b->dland.n), kMaxLandIds);
    return RK_OK;
}

// Room for kQresWords words per query in pinned, coherent host memory
// (written by the step's kernels).  Caller holds the context lock.
rk_status ensure_qres(rk_batch* b, uint64_t nq)
{
    rk_ctx* c = b->ds->ctx;
    RK_HIP(c, b->h_qres.reserve(std::max<uint64_t>(kQresWords * nq, kQresWords), hipHostMallocCoherent));
    return RK_OK;
}

// The landed extent of a finished run: every query that fit its segment has
// its ids at [qland, qland + count), densely packed from 0.  *short_area:
// some such query found the landing area full; *need: ids those queries
// need in all.
uint64_t landed_extent(const rk_batch* b, bool* short_area, uint64_t* need)
{
    uint64_t ext = 0, nd = 0;
    bool sh = false;
    const uint32_t* cnt = b->h_qcount();
    const uint32_t* land = b->h_qland();
    for (uint32_t i = 0; i < b->nq; ++i) {
        if (b->hq[i].trivial || cnt[i] > b->view.seg_cap || !cnt[i]) continue;
        nd += cnt[i];
        if (land[i] == kNotLanded)
            sh = true;
        else
            ext = std::max<uint64_t>(ext, (uint64_t)land[i] + cnt[i]);
    }
    if (short_area) *short_area = sh;
    if (need) *need = nd;
    return ext;
}

// The kernels' view of a filled batch.  Inputs: a device-filtered batch's
// queries and candidate lists are made on the device; a host-filtered one's
// are in HBM (resident) or read in place from the pinned staging.
void set_view(rk_batch* b, uint64_t seg_cap)
{
    rk_dataset* ds = b->ds;
    BatchDev& v = b->view;
    const bool dev_inputs = b->dfilter || b->resident;
    v.q = dev_inputs ? b->q.p : b->s_q.dev;
    v.nq = b->nq;
    v.seg_cap = (uint32_t)seg_cap;
    v.cand = dev_inputs ? b->cand.p : b->s_cand.dev;
    v.fx = ds->fx.p;
    v.fy = ds->fy.p;
    v.wbound = dev_inputs ? b->wb.p : b->s_wb.dev;
    v.rec = b->rec.p;
    v.orig = b->orig.p;
    v.seg = b->seg.p;
    v.qcount = b->qres.p;
    v.qland = b->qres.p + b->nq;
    v.qstat = b->qstat.p;
    v.dland = b->dland.p;
    v.hland = b->land_host ? b->land_host->dev : nullptr;
    v.hqres = b->h_qres.dev;
    v.dn = b->dfilter ? b->dn.p : nullptr;
    v.land_cap = b->land_cap;
    v.land_ctr = b->land_ctr.p;
    v.id_bits = ds->id_bits;
}

// The device filter's inputs: in HBM (resident) or the pinned staging.
const uint32_t* dfilter_q(const rk_batch* b) { return b->resident ? b->dq.p : b->s_dq.dev; }
const uint32_t* dfilter_k(const rk_batch* b) { return b->resident ? b->dk.p : b->s_dk.dev; }

// First landing area of a batch: room for 2,048 ids per query (a run that
// needs more grows it and is repeated, rk_batch_wait).
uint64_t initial_landing(uint64_t nq) { return std::max<uint64_t>(nq * 2048, 1u << 16); }

uint32_t pow2_ceil(uint64_t v)
{
    uint32_t p = 1;
    while (p < v && p < (1u << 30)) p <<= 1;
    return p;
}

// Workload class of a batch: device, tree depth, queries (power of two), mean
// candidates and mean k (powers of two).
std::string tune_key(const rk_batch* b)
{
    uint64_t nc = 0, k = 0, n = 0;
    for (const QDev& q : b->hq)
        if (!q.trivial) {
            nc += q.nc;
            k += q.k;
            ++n;
        }
    n = n ? n : 1;
    return device_arch(b->ds->ctx) + "/batch/L" + std::to_string(b->ds->view.n_levels) + "/q" +
           std::to_string(pow2_ceil(b->nq)) + "/c" + std::to_string(pow2_ceil(nc / n)) + "/k" +
           std::to_string(pow2_ceil(k / n));
}

} // namespace

namespace rk {

// Block shape for a filled batch: cached per workload class, else timed on
// the batch itself (batches of 64 queries and more; smaller ones keep the
// default).  Caller holds the context lock.
rk_status batch_tune(rk_batch* b)
{
    b->shape = kShape8Waves;
    const char* env = std::getenv("RK_AUTOTUNE"); // "0": default shape, "8" / "4": that shape
    if (env && env[0] == '4' && env[1] == 0) b->shape = kShape4Waves;
    if ((env && (env[0] == '0' || env[0] == '4' || env[0] == '8')) || !b->ds->nu || b->nq == 0) return RK_OK;
    const std::string key = tune_key(b);
    const int known = tuned_choice(key);
    if (known >= 0 && known < kNumShapes) {
        b->shape = known;
        return RK_OK;
    }
    if (b->nq < 64) return RK_OK;
    rk_ctx* c = b->ds->ctx;
    float best = 0.f;
    for (int shape = 0; shape < kNumShapes; ++shape) {
        launch_batch_query(b->ds->view, b->view, shape, c->stream); // warm
        RK_HIP(c, hipGetLastError());
        float t = 0.f;
        for (int r = 0; r < 3; ++r) {
            RK_HIP(c, hipEventRecord(c->ev0, c->stream));
            launch_batch_query(b->ds->view, b->view, shape, c->stream);
            RK_HIP(c, hipEventRecord(c->ev1, c->stream));
            RK_HIP(c, hipEventSynchronize(c->ev1));
            float ms = 0.f;
            RK_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
            t = r ? std::min(t, ms) : ms;
        }
        if (shape == 0 || t < best) {
            best = t;
            b->shape = shape;
        }
    }
    tuned_store(key, b->shape, b->shape == kShape4Waves ? "4 waves per query" : "8 waves per query");
    return RK_OK;
}

rk_status batch_fill(rk_batch* b, const rk_query_desc* qs, size_t nq)
{
    rk_dataset* ds = b->ds;
    rk_ctx* c = ds->ctx;
    if (b->pending) return fail(c, RK_ERR_INVALID_ARG, "the batch's last run has not been waited for");
    if (nq > (1u << 24)) return fail(c, RK_ERR_INVALID_ARG, "too many queries in one batch");
    b->nq = (uint32_t)nq;
    b->ran = false;
    b->dfilter = false; // a batch object may alternate between the two kinds of fill
    b->cand_overflow = false;
    b->hq.assign(nq, QDev{});
    std::vector<uint32_t> all;
    std::vector<double2> wb(nq);
    std::vector<uint32_t> cl;
    // influence-zone filter for queries without a candidate list, on the
    // host's cores
    std::vector<uint32_t> fq, fk, fslot(nq, UINT32_MAX);
    for (size_t i = 0; i < nq; ++i) {
        const rk_query_desc& d = qs[i];
        if (d.q >= ds->nf) return fail(c, RK_ERR_INVALID_ARG, "query index q must be < number of facilities");
        if (d.k == 0) return fail(c, RK_ERR_INVALID_ARG, "k must be a positive integer");
        if (d.cand == nullptr && d.nc == 0) {
            fslot[i] = (uint32_t)fq.size();
            fq.push_back(d.q);
            fk.push_back(d.k);
        }
    }
    std::vector<std::vector<uint32_t>> fc;
    const double f0 = now_s();
    if (!fq.empty()) filter_many(*ds->index, fq.data(), fk.data(), fq.size(), fc);
    const double filt = now_s() - f0;
    for (size_t i = 0; i < nq; ++i) {
        const rk_query_desc& d = qs[i];
        const rk_status s = fslot[i] != UINT32_MAX
                                ? clean_candidates(ds, d.q, fc[fslot[i]].data(), fc[fslot[i]].size(), cl)
                                : clean_candidates(ds, d.q, d.cand, d.nc, cl);
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
    b->filter_s = filt;
    b->total_cand = all.size();

    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    uint64_t seg_cap = 0;
    if (b->cap_nq) { // a fixed layout: nothing grows
        if (nq > b->cap_nq || all.size() > b->cap_nq * b->cap_cand)
            return fail(c, RK_ERR_ADMISSION, "the step exceeds the batch's fixed layout (its memory budget)");
        seg_cap = b->cap_seg;
    } else {
        RK_HIP(c, b->rec.reserve(all.size()));
        RK_HIP(c, b->orig.reserve(all.size()));
        RK_HIP(c, b->qres.reserve(2 * (uint64_t)nq));
        RK_HIP(c, b->qstat.reserve((uint64_t)nq * kQueryStats));
        // result segment per query: the whole budget split evenly, never more than the users
        seg_cap = std::max<uint64_t>(1, std::min<uint64_t>(ds->nu, nq ? pool_budget() / nq : 1));
        seg_cap = std::max<uint64_t>(seg_cap, std::min<uint64_t>(ds->nu, b->seg_min));
        seg_cap = std::min<uint64_t>(seg_cap, 0x7fffffffull);
        RK_HIP(c, b->seg.reserve(std::max<uint64_t>(1, nq * seg_cap)));
        const rk_status ls = ensure_landing(b, initial_landing(nq));
        if (ls != RK_OK) return ls;
    }
    {
        const rk_status ls = ensure_qres(b, nq);
        if (ls != RK_OK) return ls;
    }
    // the inputs into pinned staging (every previous run of this batch is
    // done: fills refuse a pending batch)
    RK_HIP(c, b->s_q.reserve(std::max<size_t>(nq, 1), hipHostMallocCoherent));
    RK_HIP(c, b->s_wb.reserve(std::max<size_t>(nq, 1), hipHostMallocCoherent));
    RK_HIP(c, b->s_cand.reserve(std::max<size_t>(all.size(), 1), hipHostMallocCoherent));
    if (nq) {
        std::memcpy(b->s_q.p, b->hq.data(), nq * sizeof(QDev));
        std::memcpy(b->s_wb.p, wb.data(), nq * sizeof(double2));
    }
    if (!all.empty()) std::memcpy(b->s_cand.p, all.data(), all.size() * sizeof(uint32_t));
    if (b->resident) { // to HBM, once (a fixed layout is never resident)
        RK_HIP(c, b->q.reserve(nq));
        RK_HIP(c, b->wb.reserve(nq));
        RK_HIP(c, b->cand.reserve(all.size()));
        if (nq) RK_HIP(c, hipMemcpyAsync(b->q.p, b->s_q.p, nq * sizeof(QDev), hipMemcpyHostToDevice, st));
        if (nq) RK_HIP(c, hipMemcpyAsync(b->wb.p, b->s_wb.p, nq * sizeof(double2), hipMemcpyHostToDevice, st));
        if (!all.empty())
            RK_HIP(c, hipMemcpyAsync(b->cand.p, b->s_cand.p, all.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                     st));
        RK_HIP(c, hipStreamSynchronize(st));
    }
    set_view(b, seg_cap);
    return RK_OK;
}

void batch_launch(rk_batch* b, hipStream_t st)
{
    rk_dataset* ds = b->ds;
    if (b->dfilter) {
        launch_dfilter(ds->fgrid, dfilter_q(b), dfilter_k(b), b->nq, b->dlists, b->cand.p, b->cand_cap, b->dn.p, st);
        launch_dfilter_setup(ds->fgrid, dfilter_q(b), dfilter_k(b), b->nq, b->dn.p, b->cand_cap, (uint32_t)ds->nu,
                             b->q.p, b->wb.p, st);
    }
    launch_batch_query(ds->view, b->view, b->shape, st);
}

rk_status batch_fill_device(rk_batch* b, const uint32_t* q, const uint32_t* k, size_t nq)
{
    rk_dataset* ds = b->ds;
    rk_ctx* c = ds->ctx;
    if (b->pending) return fail(c, RK_ERR_INVALID_ARG, "the batch's last run has not been waited for");
    if (nq > (1u << 24)) return fail(c, RK_ERR_INVALID_ARG, "too many queries in one batch");
    for (size_t i = 0; i < nq; ++i) {
        if (q[i] >= ds->nf) return fail(c, RK_ERR_INVALID_ARG, "query index q must be < number of facilities");
        if (k[i] == 0) return fail(c, RK_ERR_INVALID_ARG, "k must be a positive integer");
        if (k[i] > kDfMaxK) return fail(c, RK_ERR_INVALID_ARG, "the device filter takes k <= 64");
    }
    b->nq = (uint32_t)nq;
    b->ran = false;
    b->dfilter = true;
    b->filter_s = 0.0;
    b->total_cand = 0;
    b->hq.assign(nq, QDev{});
    for (size_t i = 0; i < nq; ++i) b->hq[i].k = k[i]; // the rest comes from each run
    b->dlists = dfilter_lists(k, nq);
    b->h_dn.assign(nq, 0);

    // candidate segment per query: every facility at most, 4096 slots at most
    // (the fixed layout's slots at most)
    b->cand_cap = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ds->nf, dfilter_slots()));
    if (b->cap_nq) b->cand_cap = std::min(b->cand_cap, b->cap_cand);
    const uint64_t slots = std::max<uint64_t>(1, (uint64_t)nq * b->cand_cap);
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    uint64_t seg_cap = 0;
    if (b->cap_nq) { // a fixed layout: nothing grows
        if (nq > b->cap_nq)
            return fail(c, RK_ERR_ADMISSION, "the step exceeds the batch's fixed layout (its memory budget)");
        seg_cap = b->cap_seg;
    } else {
        RK_HIP(c, b->dq.reserve(nq));
        RK_HIP(c, b->dk.reserve(nq));
        RK_HIP(c, b->dn.reserve(nq));
        RK_HIP(c, b->q.reserve(nq));
        RK_HIP(c, b->wb.reserve(nq));
        RK_HIP(c, b->cand.reserve(slots));
        RK_HIP(c, b->rec.reserve(slots));
        RK_HIP(c, b->orig.reserve(slots));
        RK_HIP(c, b->qres.reserve(2 * (uint64_t)nq));
        RK_HIP(c, b->qstat.reserve((uint64_t)nq * kQueryStats));
        seg_cap = std::max<uint64_t>(1, std::min<uint64_t>(ds->nu, nq ? pool_budget() / nq : 1));
        seg_cap = std::max<uint64_t>(seg_cap, std::min<uint64_t>(ds->nu, b->seg_min));
        seg_cap = std::min<uint64_t>(seg_cap, 0x7fffffffull);
        RK_HIP(c, b->seg.reserve(std::max<uint64_t>(1, nq * seg_cap)));
        const rk_status ls = ensure_landing(b, initial_landing(nq));
        if (ls != RK_OK) return ls;
    }
    {
        const rk_status ls = ensure_qres(b, nq);
        if (ls != RK_OK) return ls;
    }
    RK_HIP(c, b->s_dq.reserve(std::max<size_t>(nq, 1), hipHostMallocCoherent));
    RK_HIP(c, b->s_dk.reserve(std::max<size_t>(nq, 1), hipHostMallocCoherent));
    if (nq) {
        std::memcpy(b->s_dq.p, q, nq * sizeof(uint32_t));
        std::memcpy(b->s_dk.p, k, nq * sizeof(uint32_t));
    }
    if (b->resident && nq) { // to HBM, once
        RK_HIP(c, hipMemcpyAsync(b->dq.p, b->s_dq.p, nq * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        RK_HIP(c, hipMemcpyAsync(b->dk.p, b->s_dk.p, nq * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        RK_HIP(c, hipStreamSynchronize(st));
    }
    set_view(b, seg_cap);
    return RK_OK;
}

uint64_t batch_fixed_bytes(uint64_t nq, uint64_t seg, uint64_t cand)
{
    const uint64_t per_query = sizeof(QDev) + sizeof(double2) + 2 * sizeof(uint32_t) + kQueryStats * sizeof(uint32_t) +
                               3 * sizeof(uint32_t);
    const uint64_t per_slot = sizeof(uint32_t) + sizeof(CandRec) + sizeof(float2);
    return nq * per_query + nq * cand * per_slot + 2 * nq * seg * sizeof(int32_t) + 2 * sizeof(unsigned long long);
}

rk_status batch_fix(rk_batch* b, uint64_t nq, uint64_t seg, uint32_t cand)
{
    rk_ctx* c = b->ds->ctx;
    if (!nq || !seg || !cand || nq * seg > kMaxLandIds) return fail(c, RK_ERR_INVALID_ARG, "an empty or oversized layout");
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    RK_HIP(c, b->q.reserve(nq));
    RK_HIP(c, b->wb.reserve(nq));
    RK_HIP(c, b->qres.reserve(2 * nq));
    RK_HIP(c, b->qstat.reserve(nq * kQueryStats));
    RK_HIP(c, b->dq.reserve(nq));
    RK_HIP(c, b->dk.reserve(nq));
    RK_HIP(c, b->dn.reserve(nq));
    RK_HIP(c, b->cand.reserve(nq * cand));
    RK_HIP(c, b->rec.reserve(nq * cand));
    RK_HIP(c, b->orig.reserve(nq * cand));
    RK_HIP(c, b->seg.reserve(nq * seg));
    // the device side holds every result that fits its segment; the host
    // side (pinned host memory, not device memory) grows with the results
    RK_HIP(c, b->dland.reserve(nq * seg));
    rk_status s = ensure_landing(b, initial_landing(nq));
    if (s == RK_OK) s = ensure_qres(b, nq);
    if (s != RK_OK) return s;
    RK_HIP(c, b->s_q.reserve(nq, hipHostMallocCoherent));
    RK_HIP(c, b->s_wb.reserve(nq, hipHostMallocCoherent));
    RK_HIP(c, b->s_dq.reserve(nq, hipHostMallocCoherent));
    RK_HIP(c, b->s_dk.reserve(nq, hipHostMallocCoherent));
    RK_HIP(c, b->s_cand.reserve(nq * cand, hipHostMallocCoherent));
    b->cap_nq = nq;
    b->cap_seg = seg;
    b->cap_cand = cand;
    return RK_OK;
}

uint64_t batch_device_bytes(const rk_batch* b)
{
    return b->q.bytes() + b->wb.bytes() + b->qres.bytes() + b->qstat.bytes() + b->dq.bytes() + b->dk.bytes() +
           b->dn.bytes() + b->cand.bytes() + b->rec.bytes() + b->orig.bytes() + b->seg.bytes() + b->dland.bytes() +
           b->land_ctr.bytes();
}

// After a device-filtered run: the per-query views (trivial flag, candidate
// count) from the words the step's kernels wrote to host memory; a
// candidate list beyond its segment is an overflow.
void views_from_host(rk_batch* b)
{
    const uint32_t* nc = b->h_qncand();
    const uint32_t* fl = b->h_qflags();
    b->cand_overflow = false;
    b->total_cand = 0;
    for (uint32_t i = 0; i < b->nq; ++i) {
        b->h_dn[i] = nc[i];
        b->hq[i].nc = std::min(nc[i], b->cand_cap);
        b->hq[i].trivial = (fl[i] & kQflagTrivial) ? 1u : 0u;
        b->cand_overflow |= nc[i] > b->cand_cap;
        b->total_cand += b->hq[i].nc;
    }
}

// After a device-filtered run: the per-query views (for fetch and the
// autotuner's class) and the candidate counts (overflow check).
rk_status batch_download_views(rk_batch* b)
{
    rk_ctx* c = b->ds->ctx;
    hipStream_t st = c->stream;
    if (!b->nq) return RK_OK;
    RK_HIP(c, hipMemcpyAsync(b->hq.data(), b->q.p, b->nq * sizeof(QDev), hipMemcpyDeviceToHost, st));
    RK_HIP(c, hipMemcpyAsync(b->h_dn.data(), b->dn.p, b->nq * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    RK_HIP(c, hipStreamSynchronize(st));
    b->cand_overflow = false;
    b->total_cand = 0;
    for (uint32_t i = 0; i < b->nq; ++i) {
        b->cand_overflow |= b->h_dn[i] > b->cand_cap;
        b->total_cand += b->hq[i].nc;
    }
    return RK_OK;
}

rk_status fetch_impl(rk_batch* b, int32_t* out_ids, size_t cap, uint64_t* offsets, std::vector<uint32_t>* skip);

// One run of a filled batch on its own stream: events around the device
// work, the kernels (or the captured graph), the copy of the counts and
// landing offsets (inside the graph, or after the kernels), and the done
// event.  Caller holds the context lock.  On failure whatever reached the
// stream is finished before returning.
rk_status run_on_stream(rk_batch* b)
{
    rk_ctx* c = b->ds->ctx;
    const bool direct = !b->exec;
    // each object is created on its own, so one that failed is retried
    if (!b->ev_start) RK_HIP(c, hipEventCreate(&b->ev_start));
    if (!b->ev_end) RK_HIP(c, hipEventCreate(&b->ev_end));
    if (!b->ev_done) RK_HIP(c, hipEventCreate(&b->ev_done));
    if (!b->stream) RK_HIP(c, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)); // default priority
    // The fills finished on the context stream before they returned, and
    // fetch runs only after the wait, so the batch's stream needs no ordering
    // against the context stream.
    hipStream_t st = b->stream;
    hipError_t e = hipEventRecord(b->ev_start, st);
    const char* what = "hipEventRecord";
    if (e == hipSuccess) {
        if (direct) {
            batch_launch(b, st);
            e = hipGetLastError();
            what = "kernel launch";
        } else {
            e = hipGraphLaunch(b->exec, st);
            what = "hipGraphLaunch";
        }
    }
    if (e == hipSuccess) e = hipEventRecord(b->ev_end, st);
    if (e == hipSuccess) e = hipEventRecord(b->ev_done, st);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        return fail_hip(c, e, what, __LINE__);
    }
    return RK_OK;
}

// Copies of (dst, src, n) id ranges, split over the context's host threads
// when large (one thread copies ~10 GB/s into touched pages, less into fresh
// ones, whose first-touch faults the threads share; a step's results are
// megabytes).  Caller holds the context lock.
void copy_ranges(rk_ctx* c, const std::vector<IdRange>& r, uint64_t total)
{
    constexpr uint64_t kPerThread = 1ull << 16; // ids
    if (total < 2 * kPerThread) {
        for (const IdRange& x : r) std::memcpy(x.dst, x.src, x.n * sizeof(int32_t));
        return;
    }
    Gang& gang = host_gang(c);
    const unsigned t = (unsigned)std::min<uint64_t>(gang.size(), total / kPerThread);
    // part k < t copies the ids [k, k+1) * total / t of the concatenated ranges
    gang.run([&](unsigned k) {
        if (k >= t) return;
        const uint64_t lo = total * k / t, hi = total * (k + 1) / t;
        uint64_t at = 0;
        for (const IdRange& x : r) {
            const uint64_t a = std::max(lo, at), e = std::min(hi, at + x.n);
            if (a < e) std::memcpy(x.dst + (a - at), x.src + (a - at), (e - a) * sizeof(int32_t));
            at += x.n;
            if (at >= hi) break;
        }
    });
}

} // namespace rk

extern "C" {

rk_status rk_batch_prepare_device(rk_dataset* ds, const uint32_t* q, const uint32_t* k, size_t nq, rk_batch** out)
{
    if (!ds || !out || (nq && (!q || !k))) return RK_ERR_INVALID_ARG;
    *out = nullptr;
    std::unique_ptr<rk_batch> b(new rk_batch());
    b->ds = ds;
    b->resident = true;
    rk_status s = batch_fill_device(b.get(), q, k, nq);
    if (s != RK_OK) return s;
    if (ds->nu && nq) {
        rk_ctx* c = ds->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        // one filter pass first: the autotuner times the traversal on these lists
        launch_dfilter(ds->fgrid, b->dq.p, b->dk.p, b->nq, b->dlists, b->cand.p, b->cand_cap, b->dn.p, c->stream);
        launch_dfilter_setup(ds->fgrid, b->dq.p, b->dk.p, b->nq, b->dn.p, b->cand_cap, (uint32_t)ds->nu, b->q.p,
                             b->wb.p, c->stream);
        RK_HIP(c, hipGetLastError());
        s = batch_download_views(b.get());
        if (s == RK_OK) s = batch_tune(b.get());
        if (s == RK_OK) s = build_graph(b.get());
        if (s != RK_OK) return s;
    }
    *out = b.release();
    return RK_OK;
}

rk_status rk_batch_prepare(rk_dataset* ds, const rk_query_desc* qs, size_t nq, rk_batch** out)
{
    if (!ds || !out || (nq && !qs)) return RK_ERR_INVALID_ARG;
    *out = nullptr;
    std::unique_ptr<rk_batch> b(new rk_batch());
    b->ds = ds;
    b->resident = true;
    rk_status s = batch_fill(b.get(), qs, nq);
    if (s != RK_OK) return s;
    if (ds->nu && nq) {
        std::lock_guard<std::recursive_mutex> lk(ds->ctx->mu);
        s = batch_tune(b.get());
        if (s == RK_OK) s = build_graph(b.get());
        if (s != RK_OK) return s;
    }
    *out = b.release();
    return RK_OK;
}

rk_status rk_batch_enqueue(rk_batch* b)
{
    if (!b) return RK_ERR_INVALID_ARG;
    rk_ctx* c = b->ds->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    if (b->pending) return fail(c, RK_ERR_INVALID_ARG, "the batch's last run has not been waited for");
    // results of an earlier run are void from here on; the batch is pending
    // only once everything of this run is on its stream, so a failed call
    // leaves it free for the next fill or enqueue
    b->ran = false;
    b->overflow = false;
    b->cand_overflow = false;
    const bool direct = !b->exec && b->direct && b->ds->nu && b->nq;
    if (!b->exec && !direct) {
        b->ran = true;
        b->pending = true;
        b->pending_empty = true;
        return RK_OK;
    }
    const rk_status s = rk::run_on_stream(b);
    if (s != RK_OK) return s;
    b->pending = true;
    b->pending_empty = false;
    return RK_OK;
}

rk_status rk_batch_wait(rk_batch* b, double* device_seconds)
{
    if (!b) return RK_ERR_INVALID_ARG;
    rk_ctx* c = b->ds->ctx;
    if (!b->pending) return fail(c, RK_ERR_INVALID_ARG, "no run of this batch was enqueued");
    if (device_seconds) *device_seconds = 0.0;
    if (!b->pending_empty) {
        // waited for without the context lock: other batches may enqueue meanwhile
        const hipError_t e = hipEventSynchronize(b->ev_done);
        if (e != hipSuccess) {
            std::lock_guard<std::recursive_mutex> lk(c->mu);
            b->pending = false;
            return fail_hip(c, e, "hipEventSynchronize", __LINE__);
        }
    }
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    b->pending = false;
    if (b->pending_empty) return RK_OK;
    RK_HIP(c, hipSetDevice(c->device));
    float ms = 0.f;
    RK_HIP(c, hipEventElapsedTime(&ms, b->ev_start, b->ev_end));
    double dev_s = ms * 1e-3;
    for (int attempt = 0;; ++attempt) {
        if (b->dfilter) views_from_host(b);
        // every query that fit its segment has its sorted ids in the landing
        // area, unless the area was too small: then it grows to 1.5x what the
        // run needed and the run is repeated (once; a repeat on an area of
        // the exact size cannot fall short)
        uint64_t need = 0;
        bool short_area = false;
        (void)landed_extent(b, &short_area, &need);
        if (!short_area) break;
        if (attempt > 0) return fail(c, RK_ERR_INTERNAL, "a repeated run fell short of its landing area");
        // (a fixed layout's device side holds every result already: only
        // its host side grows)
        rk_status s = ensure_landing(b, b->cap_nq ? std::min<uint64_t>(need + need / 2, b->dland.n) : need + need / 2);
        if (s != RK_OK) return s;
        set_view(b, b->view.seg_cap);
        if (b->exec) { // the graph holds the old area's addresses: capture it again
            s = recapture(b);
            if (s != RK_OK) return s;
        }
        s = run_on_stream(b);
        if (s != RK_OK) return s;
        RK_HIP(c, hipEventSynchronize(b->ev_done));
        RK_HIP(c, hipEventElapsedTime(&ms, b->ev_start, b->ev_end));
        dev_s += ms * 1e-3;
    }
    if (device_seconds) *device_seconds = dev_s;
    // both overflow checks before either is reported: fetch refuses a run
    // with either kind (a truncated candidate list gives too many results, a
    // truncated segment too few)
    for (uint32_t i = 0; i < b->nq; ++i) b->overflow |= b->h_qcount()[i] > b->view.seg_cap;
    b->ran = true;
    if (b->cand_overflow)
        return fail(c, RK_ERR_CAPACITY, "a query's device candidate list exceeds its segment; filter on the host");
    if (b->overflow) return fail(c, RK_ERR_CAPACITY, "a query's result exceeds its segment; split the batch");
    return RK_OK;
}

rk_status rk_batch_run(rk_batch* b, double* device_seconds)
{
    const rk_status s = rk_batch_enqueue(b);
    if (s != RK_OK) return s;
    return rk_batch_wait(b, device_seconds);
}

rk_status rk_batch_fetch(rk_batch* b, int32_t* out_ids, size_t cap, uint64_t* offsets)
{
    if (!b || !offsets || (cap && !out_ids)) return RK_ERR_INVALID_ARG;
    return rk::fetch_impl(b, out_ids, cap, offsets, nullptr);
}

rk_status rk_batch_results(rk_batch* b, const int32_t** ids, uint64_t* n)
{
    if (!b || (b->nq && (!ids || !n))) return RK_ERR_INVALID_ARG;
    rk_ctx* c = b->ds->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    std::vector<IdRange> ranges;
    const rk_status s = rk::result_ranges(b, ranges, nullptr);
    if (s != RK_OK) return s;
    for (uint32_t i = 0; i < b->nq; ++i) {
        ids[i] = ranges[i].src;
        n[i] = ranges[i].n;
    }
    return RK_OK;
}

} // extern "C"

namespace rk {

// Where the last run's results are: per query a range of the landing area
// (or of the dataset's sorted id list, for queries every user qualifies
// for).  With skip == nullptr every query must have fit its segments;
// otherwise the queries that did not get empty ranges and are listed in
// *skip.  Caller holds the context lock.
rk_status result_ranges(rk_batch* b, std::vector<IdRange>& ranges, std::vector<uint32_t>* skip)
{
    rk_dataset* ds = b->ds;
    rk_ctx* c = ds->ctx;
    if (!b->ran) return fail(c, RK_ERR_INVALID_ARG, "rk_batch_run has not been called");
    if (b->pending) return fail(c, RK_ERR_INVALID_ARG, "rk_batch_wait has not been called");
    if (b->cand_overflow && !skip)
        return fail(c, RK_ERR_CAPACITY, "the last run overflowed a device candidate segment; filter on the host");
    if (b->overflow && !skip)
        return fail(c, RK_ERR_CAPACITY, "the last run overflowed a result segment; split the batch");
    const uint32_t nq = b->nq;
    const uint32_t* cnt = b->h_qcount();
    const uint32_t* land = b->h_qland();
    if (skip) skip->clear();
    ranges.assign(nq, IdRange{nullptr, nullptr, 0});
    for (uint32_t i = 0; i < nq; ++i) {
        const QDev& h = b->hq[i];
        // a truncated device candidate list first: its view may say trivial
        if (cnt[i] > b->view.seg_cap || (b->dfilter && b->h_dn[i] > b->cand_cap)) {
            skip->push_back(i); // (without skip the overflow was refused above)
            continue;
        }
        if (h.trivial) {
            if (h.k > h.nc) ranges[i] = {nullptr, ds->all_ids_sorted.data(), ds->nu};
            continue;
        }
        if (!cnt[i]) continue;
        if (land[i] == kNotLanded || (uint64_t)land[i] + cnt[i] > b->land_cap)
            return fail(c, RK_ERR_INTERNAL, "a result is missing from the landing area");
        ranges[i] = {nullptr, b->land_host->p + land[i], cnt[i]};
    }
    return RK_OK;
}

// The results of the last run, copied from the landing area (see
// result_ranges for skip).
rk_status fetch_impl(rk_batch* b, int32_t* out_ids, size_t cap, uint64_t* offsets, std::vector<uint32_t>* skip)
{
    rk_ctx* c = b->ds->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    std::vector<IdRange> ranges;
    const rk_status s = result_ranges(b, ranges, skip);
    if (s != RK_OK) return s;
    const uint32_t nq = b->nq;
    uint64_t total = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        offsets[i] = total;
        ranges[i].dst = out_ids + total;
        total += ranges[i].n;
    }
    offsets[nq] = total;
    if (total > cap) return fail(c, RK_ERR_CAPACITY, "output buffer too small; offsets[nq] holds the size needed");
    copy_ranges(c, ranges, total);
    return RK_OK;
}

rk_status batch_fetch_partial(rk_batch* b, std::vector<int32_t>& ids, std::vector<uint64_t>& off,
                              std::vector<uint32_t>& redo)
{
    off.assign(b->nq + 1, 0);
    ids.resize(std::max<size_t>(ids.size(), 1));
    while (true) {
        const rk_status s = fetch_impl(b, ids.data(), ids.size(), off.data(), &redo);
        if (s != RK_ERR_CAPACITY || off[b->nq] <= ids.size()) return s;
        ids.resize(off[b->nq]); // the size needed is known now
    }
}

rk_status batch_collect(rk_batch* b, rk_batch* r, const rk_query_desc* qd, size_t nq, std::vector<int32_t>& ids,
                        std::vector<uint64_t>& off)
{
    rk_dataset* ds = b->ds;
    std::vector<uint32_t> redo;
    rk_status st = batch_fetch_partial(b, ids, off, redo);
    if (st != RK_OK || redo.empty()) return st;
    const bool dfilter = b->dfilter;
    std::vector<rk_query_desc> rq(redo.size());
    std::vector<std::vector<uint32_t>> cands(redo.size());
    uint64_t need = 0;
    bool host = !dfilter;
    for (size_t j = 0; j < redo.size(); ++j) {
        const uint32_t i = redo[j];
        rq[j] = qd[i];
        need = std::max<uint64_t>(need, b->h_qcount()[i]);
        host |= dfilter && b->h_dn[i] > b->cand_cap;
    }
    if (host) // candidate lists from the host filter for every re-run query without one
        for (size_t j = 0; j < rq.size(); ++j)
            if (!rq[j].cand) {
                ds->index->filter(rq[j].q, rq[j].k, cands[j]);
                rq[j].cand = cands[j].data();
                rq[j].nc = cands[j].size();
            }
    std::vector<int32_t> rids;
    std::vector<uint64_t> roff;
    r->ds = ds;
    r->direct = true;
    // the re-run queries [j0, j1) in one run; segments hold the counts the
    // step measured and host lists never overflow, so nothing is left over
    auto rerun = [&](size_t j0, size_t j1, uint64_t seg_need) -> rk_status {
        if (r->cap_nq && seg_need > r->cap_seg)
            return fail(ds->ctx, RK_ERR_ADMISSION, "a result larger than the memory budget's re-run share");
        r->seg_min = seg_need;
        rk_status s;
        if (host) {
            s = batch_fill(r, rq.data() + j0, j1 - j0);
        } else {
            std::vector<uint32_t> q(j1 - j0), k(j1 - j0);
            for (size_t j = j0; j < j1; ++j) {
                q[j - j0] = rq[j].q;
                k[j - j0] = rq[j].k;
            }
            s = batch_fill_device(r, q.data(), k.data(), q.size());
        }
        if (s == RK_OK) s = rk_batch_run(r, nullptr);
        if (s != RK_OK && s != RK_ERR_CAPACITY) return s;
        std::vector<int32_t> ids1;
        std::vector<uint64_t> off1;
        std::vector<uint32_t> again;
        s = batch_fetch_partial(r, ids1, off1, again);
        if (s == RK_OK && !again.empty())
            return fail(ds->ctx, RK_ERR_INTERNAL, "a re-run query outgrew segments sized to its count");
        if (s != RK_OK) return s;
        if (roff.empty()) roff.push_back(0);
        for (size_t j = 0; j < j1 - j0; ++j) {
            rids.insert(rids.end(), ids1.begin() + off1[j], ids1.begin() + off1[j + 1]);
            roff.push_back(rids.size());
        }
        return RK_OK;
    };
    if (r->cap_nq) { // a fixed layout: one query at a time, within its segment
        for (size_t j = 0; j < rq.size() && st == RK_OK; ++j)
            st = rerun(j, j + 1, b->h_qcount()[redo[j]]);
    } else {
        st = rerun(0, rq.size(), need);
    }
    if (st != RK_OK) return st;
    std::vector<int32_t> all;
    all.reserve(off[nq] + roff[rq.size()]);
    std::vector<uint64_t> noff(nq + 1, 0);
    for (size_t i = 0, j = 0; i < nq; ++i) {
        noff[i] = all.size();
        if (j < redo.size() && redo[j] == i) {
            all.insert(all.end(), rids.begin() + roff[j], rids.begin() + roff[j + 1]);
            ++j;
        } else {
            all.insert(all.end(), ids.begin() + off[i], ids.begin() + off[i + 1]);
        }
    }
    noff[nq] = all.size();
    ids.swap(all);
    off.swap(noff);
    return RK_OK;
}

} // namespace rk

extern "C" {

rk_status rk_batch_stats(const rk_batch* b, uint64_t* items, size_t n_items_cap, size_t* n_levels, uint64_t* results,
                         uint64_t* rechecks)
{
    if (!b || !b->ran || b->pending) return RK_ERR_INVALID_ARG;
    rk_ctx* c = b->ds->ctx;
    const size_t nl = b->ds->view.n_levels;
    if (n_levels) *n_levels = nl;
    std::vector<uint32_t> st;
    const bool dev_ran = b->exec || (b->direct && b->ds->nu && b->nq);
    if (dev_ran) {
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        RK_HIP(c, hipSetDevice(c->device));
        st.resize((size_t)b->nq * kQueryStats);
        RK_HIP(c, hipMemcpyAsync(st.data(), b->qstat.p, st.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 c->stream));
        RK_HIP(c, hipStreamSynchronize(c->stream));
    }
    uint64_t lv[kMaxLevels] = {}, rech = 0, res = 0;
    for (uint32_t i = 0; dev_ran && i < b->nq; ++i) {
        const uint32_t* s = st.data() + (size_t)i * kQueryStats;
        rech += s[0];
        for (size_t l = 0; l < nl; ++l) lv[l] += s[1 + l];
        res += b->h_qcount()[i];
    }
    for (size_t l = 0; l < nl && l < n_items_cap && items; ++l) items[l] = lv[l];
    if (results) *results = res;
    if (rechecks) *rechecks = rech;
    return RK_OK;
}

rk_status rk_batch_config(const rk_batch* b, uint32_t* waves_per_query)
{
    if (!b || !waves_per_query) return RK_ERR_INVALID_ARG;
    *waves_per_query = b->shape == kShape4Waves ? 4 : 8;
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

} // extern "C"

namespace {

// Queries without candidate lists and with k <= 64 are filtered on the
// device (inside the batch's graph) unless RK_FILTER=host; a query whose
// device candidate list outgrew its segment runs again through the host
// filter (batch_collect).
bool device_filterable(const rk_query_desc* qs, size_t nq)
{
    const char* env = std::getenv("RK_FILTER");
    if (env && std::strcmp(env, "host") == 0) return false;
    if (!nq) return false;
    for (size_t i = 0; i < nq; ++i)
        if (qs[i].cand || qs[i].nc || qs[i].k == 0 || qs[i].k > kDfMaxK) return false;
    return true;
}

} // namespace

extern "C" rk_status rk_query_batch(rk_dataset* ds, const rk_query_desc* qs, size_t nq, int32_t* out_ids, size_t cap,
                                    uint64_t* offsets, rk_timing* t)
{
    if (!ds || !offsets || (nq && !qs)) return RK_ERR_INVALID_ARG;
    const double t0 = now_s();
    if (t) *t = rk_timing{};
    // one run: the kernels are launched directly (a graph pays for its
    // capture only when replayed); host-filtered batches take the
    // autotuner's shape for their class (timed once per class), device-
    // filtered ones the default, as their lists exist only inside the run.
    // The dataset's batch object for the filter side is refilled in place.
    const bool on_device = device_filterable(qs, nq);
    const int side = on_device ? 0 : 1;
    std::unique_lock<std::mutex> own(ds->qb_mu[side], std::try_to_lock);
    std::unique_ptr<rk_batch> mine;
    rk_batch* b = nullptr;
    if (own.owns_lock()) {
        if (!ds->qb[side]) ds->qb[side].reset(new rk_batch());
        b = ds->qb[side].get();
    } else {
        mine.reset(new rk_batch());
        b = mine.get();
    }
    b->ds = ds;
    b->direct = true;
    rk_status s;
    if (on_device) {
        std::vector<uint32_t> hq(nq), hk(nq);
        for (size_t i = 0; i < nq; ++i) {
            hq[i] = qs[i].q;
            hk[i] = qs[i].k;
        }
        s = batch_fill_device(b, hq.data(), hk.data(), nq);
    } else {
        s = batch_fill(b, qs, nq);
        if (s == RK_OK) {
            std::lock_guard<std::recursive_mutex> lk(ds->ctx->mu);
            s = batch_tune(b);
        }
    }
    if (s != RK_OK) return s;
    const double t1 = now_s();
    double dev = 0.0;
    s = rk_batch_run(b, &dev);
    if (s != RK_OK && s != RK_ERR_CAPACITY) return s;
    const double t2 = now_s();
    if (s == RK_OK) {
        s = rk_batch_fetch(b, out_ids, cap, offsets);
    } else {
        // a result (or device candidate list) beyond its segment: the others
        // come from this run, those queries run again (batch_collect)
        std::vector<int32_t> ids;
        std::vector<uint64_t> off;
        rk_batch redo;
        redo.ds = ds;
        s = batch_collect(b, &redo, qs, nq, ids, off);
        if (s == RK_OK) {
            std::copy(off.begin(), off.end(), offsets);
            if (off[nq] > cap)
                s = fail(ds->ctx, RK_ERR_CAPACITY, "output buffer too small; offsets[nq] holds the size needed");
            else if (off[nq])
                std::memcpy(out_ids, ids.data(), off[nq] * sizeof(int32_t));
        }
    }
    if (t) {
        t->filter_s = b->filter_s;
        t->upload_s = t1 - t0 - b->filter_s;
        t->kernel_s = dev;
        t->download_s = now_s() - t2;
        t->n_candidates = b->total_cand;
        uint64_t rech = 0;
        if (rk_batch_stats(b, nullptr, 0, nullptr, nullptr, &rech) == RK_OK) t->n_fp64_rechecks = rech;
        t->total_s = now_s() - t0;
    }
    return s; // every copy out of the batch was synchronous
}
