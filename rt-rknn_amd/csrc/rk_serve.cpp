// rk_serve.cpp -- continuous-batching request loop over resident datasets.
//
// No reference counterpart: every reference program answers one query per
// process (rknn-inf-gpu/inf.cpp:631-658).  Here requests (dataset, q, k)
// arrive from any thread.  Two stages run concurrently:
//   * the host stage (one worker per host core, up to 16) admits requests;
//     a request with k <= 64 goes straight to the device stage (its filter
//     runs on the device, inside the step), any other request is filtered
//     on the host (influence-zone candidates) while the current device step
//     runs, and
//   * the device stage runs one batched step over every admitted request of
//     a dataset (device-filtered and host-filtered requests in separate
//     steps), refilling that dataset's batch object in place.
// RK_FILTER=host sends every request through the host filter.
// Admission is by memory: each request reserves an estimate of the device
// bytes it needs while in flight; a request waits until the reservation fits
// the budget, and one that could never fit is rejected at submission.
// Results do not depend on batch-mates: each query's set is computed only
// from its own candidates.
#include <algorithm>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rk_internal.hpp"
#include "rknn.h"

namespace {

struct Request {
    uint64_t ticket;
    rk_dataset* ds;
    uint32_t q, k;
    uint64_t reserve;
    double t_submit;
    bool dev;                   // filtered on the device, inside the step
    std::vector<uint32_t> cand; // host-filtered requests: filled by the host stage
};

struct Done {
    std::vector<int32_t> ids;
    double latency;
    rk_status status;
};

} // namespace

struct rk_server {
    rk_ctx* ctx = nullptr;
    rk_serve_config cfg{};
    std::mutex mu;
    std::condition_variable cv;      // wakes the host stage
    std::condition_variable cv_done; // wakes pollers (unused by the C API, kept for stop)
    std::deque<Request> waiting;     // submitted, not yet admitted
    std::deque<Request> ready;       // admitted and filtered, waiting for a device step
    std::unordered_map<uint64_t, Done> done;
    uint64_t next_ticket = 1;
    uint64_t reserved = 0;
    bool stop = false;
    std::vector<std::thread> host_threads; // filter workers
    std::thread dev_thread;
    std::condition_variable cv_dev;
    std::map<rk_dataset*, std::unique_ptr<rk_batch>> batches;  // device stage only: host-filtered steps
    std::map<rk_dataset*, std::unique_ptr<rk_batch>> dbatches; // device-filtered steps
    bool host_filter_only = false;                             // RK_FILTER=host
};

namespace {

uint64_t estimate_bytes(const rk_dataset* ds, uint32_t k, bool dev)
{
    // candidate records + indices (|C| grows roughly linearly in k; a
    // device-filtered request holds a whole candidate segment), result
    // pool entries (12 B each) for the expected result size k*N_u/N_f with
    // slack, and one tile list entry per tile of the zone's neighbourhood
    const uint64_t nf = ds->nf ? ds->nf : 1;
    const uint64_t cand = dev ? std::min<uint64_t>(nf, rk::kDfBatchCand) : 64ull * k + 256;
    uint64_t results = 8ull * k * ds->nu / nf + 4096;
    results = std::min<uint64_t>(results, ds->nu + 64);
    return cand * 28 + results * 16 + 8 * 4096 + 4096;
}

void host_stage(rk_server* s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->cv.wait(lk, [&] { return s->stop || (!s->waiting.empty() && s->reserved + s->waiting.front().reserve <= s->cfg.memory_budget); });
        if (s->stop) return;
        Request r = std::move(s->waiting.front());
        s->waiting.pop_front();
        s->reserved += r.reserve;
        if (!s->waiting.empty()) s->cv.notify_one();
        if (!r.dev) {
            lk.unlock();
            r.ds->index->filter(r.q, r.k, r.cand);
            lk.lock();
        }
        s->ready.push_back(std::move(r));
        s->cv_dev.notify_one();
    }
}

// One device step over qd.  Each dataset keeps one batch object whose
// device buffers only grow: refilled every step and launched without a
// graph, it allocates nothing once warm.  A step over its result budget is
// split by rk_query_batch.
rk_status run_step(rk_server* s, rk_dataset* ds, const std::vector<rk_query_desc>& qd, std::vector<int32_t>& ids,
                   std::vector<uint64_t>& off)
{
    std::unique_ptr<rk_batch>& b = s->batches[ds];
    if (!b) {
        b.reset(new rk_batch());
        b->ds = ds;
        b->direct = true;
    }
    rk_status st = rk::batch_fill(b.get(), qd.data(), qd.size());
    if (st == RK_OK) {
        std::lock_guard<std::recursive_mutex> lk(ds->ctx->mu);
        st = rk::batch_tune(b.get());
    }
    if (st == RK_OK) st = rk_batch_run(b.get(), nullptr);
    if (st == RK_OK) {
        uint64_t total = 0; // the result size is known once the step ran
        for (size_t i = 0; i < qd.size(); ++i) {
            const rk::QDev& h = b->hq[i];
            total += h.trivial ? (h.k > h.nc ? ds->nu : 0) : std::min(b->h_qcount.p[i], b->view.seg_cap);
        }
        ids.resize(std::max<uint64_t>(1, total));
        return rk_batch_fetch(b.get(), ids.data(), ids.size(), off.data());
    }
    if (st != RK_ERR_CAPACITY) return st;
    size_t cap = 1 << 20;
    while (true) {
        ids.resize(cap);
        st = rk_query_batch(ds, qd.data(), qd.size(), ids.data(), ids.size(), off.data(), nullptr);
        if (st == RK_ERR_CAPACITY && off[qd.size()] > cap) {
            cap = off[qd.size()];
            continue;
        }
        return st;
    }
}

// One device-filtered step over (q, k): the dataset's device batch, refilled
// in place and launched without a graph (filter, setup, traversal, sort).  A
// candidate list beyond its device segment sends the step to rk_query_batch,
// which filters it on the host.
rk_status run_step_device(rk_server* s, rk_dataset* ds, const std::vector<uint32_t>& q,
                          const std::vector<uint32_t>& k, std::vector<int32_t>& ids, std::vector<uint64_t>& off)
{
    std::unique_ptr<rk_batch>& b = s->dbatches[ds];
    if (!b) {
        b.reset(new rk_batch());
        b->ds = ds;
        b->direct = true;
    }
    // traversal shape: the default (8 waves per query); timing shapes needs
    // filtered lists, which only the step itself makes
    rk_status st = rk::batch_fill_device(b.get(), q.data(), k.data(), q.size());
    if (st == RK_OK) st = rk_batch_run(b.get(), nullptr);
    if (st == RK_OK) {
        uint64_t total = 0;
        for (size_t i = 0; i < q.size(); ++i) {
            const rk::QDev& h = b->hq[i];
            total += h.trivial ? (h.k > h.nc ? ds->nu : 0) : std::min(b->h_qcount.p[i], b->view.seg_cap);
        }
        ids.resize(std::max<uint64_t>(1, total));
        return rk_batch_fetch(b.get(), ids.data(), ids.size(), off.data());
    }
    if (st != RK_ERR_CAPACITY) return st;
    std::vector<rk_query_desc> qd(q.size());
    for (size_t i = 0; i < q.size(); ++i) qd[i] = rk_query_desc{q[i], k[i], nullptr, 0};
    size_t cap = 1 << 20;
    while (true) {
        ids.resize(cap);
        st = rk_query_batch(ds, qd.data(), qd.size(), ids.data(), ids.size(), off.data(), nullptr);
        if (st == RK_ERR_CAPACITY && off[qd.size()] > cap) {
            cap = off[qd.size()];
            continue;
        }
        return st;
    }
}

void dev_stage(rk_server* s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->cv_dev.wait(lk, [&] { return s->stop || !s->ready.empty(); });
        if (s->stop && s->ready.empty()) return;
        // let batch-mates gather briefly when the step would be small
        if (s->ready.size() < s->cfg.max_batch && s->cfg.max_wait_s > 0 && !s->stop) {
            const double deadline = s->ready.front().t_submit + s->cfg.max_wait_s;
            while (!s->stop && s->ready.size() < s->cfg.max_batch && !s->waiting.empty() && rk::now_s() < deadline) {
                s->cv_dev.wait_for(lk, std::chrono::microseconds(200));
            }
        }
        // one step: requests of the dataset (and filter side) at the head of the queue
        rk_dataset* ds = s->ready.front().ds;
        const bool dev = s->ready.front().dev;
        std::vector<Request> step;
        for (auto it = s->ready.begin(); it != s->ready.end() && step.size() < s->cfg.max_batch;) {
            if (it->ds == ds && it->dev == dev) {
                step.push_back(std::move(*it));
                it = s->ready.erase(it);
            } else {
                ++it;
            }
        }
        lk.unlock();
        std::vector<uint64_t> off(step.size() + 1);
        std::vector<int32_t> ids;
        rk_status st;
        if (dev) {
            std::vector<uint32_t> q(step.size()), k(step.size());
            for (size_t i = 0; i < step.size(); ++i) {
                q[i] = step[i].q;
                k[i] = step[i].k;
            }
            st = run_step_device(s, ds, q, k, ids, off);
        } else {
            std::vector<rk_query_desc> qd(step.size());
            for (size_t i = 0; i < step.size(); ++i) {
                qd[i].q = step[i].q;
                qd[i].k = step[i].k;
                qd[i].cand = step[i].cand.data();
                qd[i].nc = step[i].cand.size();
            }
            st = run_step(s, ds, qd, ids, off);
        }
        const double t_end = rk::now_s();
        // results are copied out before the lock is taken: submitters and
        // pollers share it
        std::vector<Done> fin(step.size());
        for (size_t i = 0; i < step.size(); ++i) {
            fin[i].status = st;
            fin[i].latency = t_end - step[i].t_submit;
            if (st == RK_OK) fin[i].ids.assign(ids.begin() + off[i], ids.begin() + off[i + 1]);
        }
        lk.lock();
        bool freed = false;
        for (size_t i = 0; i < step.size(); ++i) {
            s->done[step[i].ticket] = std::move(fin[i]);
            s->reserved -= step[i].reserve;
            freed |= step[i].reserve > 0;
        }
        if (freed && !s->waiting.empty()) s->cv.notify_all();
        s->cv_done.notify_all();
    }
}

} // namespace

extern "C" {

rk_status rk_serve_start(rk_ctx* ctx, const rk_serve_config* cfg, rk_server** out)
{
    if (!ctx || !out) return RK_ERR_INVALID_ARG;
    rk_server* s = new rk_server();
    s->ctx = ctx;
    s->cfg.max_batch = 256;
    s->cfg.memory_budget = 4ull << 30;
    s->cfg.max_wait_s = 0.0005;
    if (cfg) {
        if (cfg->max_batch) s->cfg.max_batch = cfg->max_batch;
        if (cfg->memory_budget) s->cfg.memory_budget = cfg->memory_budget;
        if (cfg->max_wait_s >= 0) s->cfg.max_wait_s = cfg->max_wait_s;
    }
    const char* env = std::getenv("RK_FILTER");
    s->host_filter_only = env && std::strcmp(env, "host") == 0;
    const unsigned nf = rk::filter_threads(1u << 20);
    for (unsigned i = 0; i < nf; ++i) s->host_threads.emplace_back(host_stage, s);
    s->dev_thread = std::thread(dev_stage, s);
    *out = s;
    return RK_OK;
}

rk_status rk_serve_submit(rk_server* s, rk_dataset* ds, uint32_t q, uint32_t k, uint64_t* ticket)
{
    if (!s || !ds || !ticket) return RK_ERR_INVALID_ARG;
    if (ds->ctx != s->ctx) return rk::fail(s->ctx, RK_ERR_INVALID_ARG, "dataset belongs to another context");
    if (q >= ds->nf || k == 0) return rk::fail(s->ctx, RK_ERR_INVALID_ARG, "q must be < N_f and k >= 1");
    Request r;
    r.ds = ds;
    r.q = q;
    r.k = k;
    r.dev = !s->host_filter_only && k <= rk::kDfMaxK;
    r.reserve = estimate_bytes(ds, k, r.dev);
    if (r.reserve > s->cfg.memory_budget)
        return rk::fail(s->ctx, RK_ERR_ADMISSION, "request can never fit the serving memory budget");
    r.t_submit = rk::now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return RK_ERR_INVALID_ARG;
    r.ticket = s->next_ticket++;
    *ticket = r.ticket;
    if (r.dev && s->waiting.empty() && s->reserved + r.reserve <= s->cfg.memory_budget) {
        // nothing to filter on the host: admitted here, straight to the device stage
        s->reserved += r.reserve;
        s->ready.push_back(std::move(r));
        s->cv_dev.notify_one();
        return RK_OK;
    }
    s->waiting.push_back(std::move(r));
    s->cv.notify_one();
    return RK_OK;
}

rk_status rk_serve_poll(rk_server* s, uint64_t ticket, int32_t* out_ids, size_t cap, size_t* n_out, double* latency_s)
{
    if (!s || !n_out || (cap && !out_ids)) return RK_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->done.find(ticket);
    if (it == s->done.end()) {
        if (ticket == 0 || ticket >= s->next_ticket) return RK_ERR_INVALID_ARG;
        return RK_ERR_NOT_READY;
    }
    Done& d = it->second;
    *n_out = d.ids.size();
    if (latency_s) *latency_s = d.latency;
    if (d.status != RK_OK) {
        rk_status st = d.status;
        s->done.erase(it);
        return st;
    }
    if (d.ids.size() > cap) return RK_ERR_CAPACITY; // keep the result for a larger buffer
    if (!d.ids.empty()) std::memcpy(out_ids, d.ids.data(), d.ids.size() * sizeof(int32_t));
    s->done.erase(it);
    return RK_OK;
}

rk_status rk_serve_stop(rk_server* s)
{
    if (!s) return RK_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    s->cv_dev.notify_all();
    for (auto& t : s->host_threads)
        if (t.joinable()) t.join();
    if (s->dev_thread.joinable()) s->dev_thread.join();
    delete s;
    return RK_OK;
}

} // extern "C"
