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
// Admission is by memory: the server's device memory is three batch objects
// (two step slots that alternate, one for re-runs), shared by every dataset
// and laid out once from memory_budget (45% per step slot, 10% for re-runs);
// a step takes only the requests its slot's layout holds (queries, candidate
// slots, result segments), so the server never allocates beyond the budget,
// and a request is rejected at submission when no step could hold it.
// Results do not depend on batch-mates: each query's set is computed only
// from its own candidates.
#include <algorithm>
#include <condition_variable>
#include <cstdio>
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
    double t_submit;
    bool dev;                   // filtered on the device, inside the step
    std::vector<uint32_t> cand; // host-filtered requests: filled by the host stage
};

// A finished request: its n ids at ids, inside memory that keep holds (the
// step's landing area, shared by the step's requests, or a step's merged
// array when some of them ran again), copied out once, by the poller.
struct Done {
    std::shared_ptr<const void> keep;
    const int32_t* ids = nullptr;
    uint64_t n = 0;
    double latency = 0.0;
    rk_status status = RK_OK;
};

} // namespace

// Result arrays of finished steps, returned by the last poller instead of
// freed: the device stage refills them with their pages still mapped (a
// fresh multi-megabyte array per step page-faults on the device stage while
// pollers unmap the old ones, and both wait on the process's memory map).
constexpr size_t kPoolKeep = 8; // arrays kept
struct IdPool {
    std::mutex mu;
    std::vector<std::vector<int32_t>*> free;
    ~IdPool()
    {
        for (auto* v : free) delete v;
    }
};

// Host landing areas of finished steps, returned by the last poller of a
// step instead of freed (pinned allocations are slow and freeing one can
// wait for the device): a batch whose area is still referenced by served
// results takes one of these for its next step.
// A keeper thread allocates areas ahead of need (a pinned allocation of a
// few megabytes takes milliseconds), so that the device stage finds kSpare
// free ones when results of earlier steps are still being polled.
constexpr size_t kAreasKeep = 32; // areas kept
constexpr size_t kSpare = 3;      // free areas the keeper maintains
struct AreaPool {
    std::mutex mu;
    std::condition_variable cv; // wakes the keeper
    bool stop = false;
    std::vector<rk::PinnedIds*> free;
    size_t largest = 0; // every new area is at least as large as the largest so far
    ~AreaPool()
    {
        for (auto* a : free) delete a;
    }
};

// The keeper: tops the free list up to kSpare areas of the largest size.
static void area_keeper(std::shared_ptr<AreaPool> pool)
{
    std::unique_lock<std::mutex> lk(pool->mu);
    while (true) {
        pool->cv.wait(lk, [&] { return pool->stop || (pool->largest && pool->free.size() < kSpare); });
        if (pool->stop) return;
        const size_t n = pool->largest;
        lk.unlock();
        rk::PinnedIds* a = new rk::PinnedIds();
        const bool ok = a->alloc(n) == hipSuccess;
        lk.lock();
        if (!ok) { // no more pinned memory: the device stage allocates (or fails) itself
            delete a;
            pool->cv.wait(lk, [&] { return pool->stop || pool->largest > n; });
            continue;
        }
        pool->free.push_back(a);
    }
}

// The device-memory layout of a server's batch objects (see the header):
// queries per step, ids per result segment and candidate slots per query of
// the step slots, and the re-run object's segment and candidate slots (one
// query at a time).
struct ServeLayout {
    bool ok = false;
    uint64_t nq = 0, seg = 0, re_seg = 0;
    uint32_t cand = 0, re_cand = 0;
};

struct rk_server {
    // first members: destroyed last, after the finished results that return to them
    std::shared_ptr<IdPool> pool = std::make_shared<IdPool>();
    std::shared_ptr<AreaPool> areas = std::make_shared<AreaPool>();
    rk_ctx* ctx = nullptr;
    rk_serve_config cfg{};
    std::mutex mu;
    std::condition_variable cv;      // wakes the host stage
    std::condition_variable cv_done; // wakes pollers (unused by the C API, kept for stop)
    std::deque<Request> waiting;     // submitted, not yet admitted
    // admitted and filtered, waiting for a device step: one queue per
    // (dataset, filtered on the device) -- the requests a step may take
    std::map<std::pair<rk_dataset*, bool>, std::deque<Request>> ready;
    size_t n_ready = 0;
    std::unordered_map<uint64_t, Done> done;
    uint64_t next_ticket = 1;
    bool stop = false;
    ServeLayout layout;
    std::vector<std::thread> host_threads; // filter workers
    std::thread dev_thread;
    std::thread keeper;                    // allocates landing areas ahead of need
    std::condition_variable cv_dev;
    // device stage only: two step slots (a step is filled and enqueued on
    // one while the previous step's run, on the other, finishes) and the
    // re-run object, for every dataset and both filter sides
    std::unique_ptr<rk_batch> slots[2];
    std::unique_ptr<rk_batch> rerun;
    bool host_filter_only = false; // RK_FILTER=host
};

namespace {

// Candidate slots per query of a step slot: a device list of k <= 64 is
// about 5k + 25 facilities (20% more than the host filter's), so 16 k + 512
// at the largest k leaves room to spare; a longer list runs again through
// the host filter.
constexpr uint32_t kServeCand = 16 * rk::kDfMaxK + 512;
constexpr uint64_t kServeSegMin = 4096, kServeSegMax = 1u << 16; // result segment ids

ServeLayout serve_layout(uint64_t budget, uint32_t max_batch)
{
    ServeLayout L;
    const uint64_t step_share = budget / 100 * 45, re_share = budget / 100 * 10;
    for (uint64_t nq = max_batch; nq >= 1; nq = nq > 3 ? nq * 3 / 4 : nq - 1) {
        const uint64_t fixed = rk::batch_fixed_bytes(nq, 0, kServeCand);
        if (fixed >= step_share) continue;
        // segments: the pool split over the step's queries, at most what the
        // share leaves (kServeSegMin at least, unless the pool is smaller)
        const uint64_t want = std::max<uint64_t>(1, std::min<uint64_t>(kServeSegMax, rk::batch_pool_ids() / nq));
        const uint64_t seg = std::min<uint64_t>(want, (step_share - fixed) / (2 * sizeof(int32_t) * nq));
        if (seg >= std::min(kServeSegMin, want)) {
            L.nq = nq;
            L.seg = seg;
            L.cand = kServeCand;
            break;
        }
    }
    // the re-run object: one query, half its share for candidates, half for
    // the segment and its landing twin
    const uint64_t per_slot = sizeof(uint32_t) + sizeof(rk::CandRec) + sizeof(float2);
    const uint64_t half = re_share > 4096 ? (re_share - 4096) / 2 : 0;
    L.re_cand = (uint32_t)std::min<uint64_t>(half / per_slot, 1u << 30);
    L.re_seg = std::min<uint64_t>(half / (2 * sizeof(int32_t)), 1ull << 30);
    L.ok = L.nq > 0 && L.re_cand > 0 && L.re_seg > 0;
    return L;
}

// caller holds s->mu
void push_ready(rk_server* s, Request&& r)
{
    s->ready[std::make_pair(r.ds, r.dev)].push_back(std::move(r));
    ++s->n_ready;
    s->cv_dev.notify_one();
}

// The queue whose head waited longest (caller holds s->mu, n_ready > 0).
std::deque<Request>& oldest_ready(rk_server* s)
{
    std::deque<Request>* best = nullptr;
    for (auto& kv : s->ready)
        if (!kv.second.empty() && (!best || kv.second.front().t_submit < best->front().t_submit)) best = &kv.second;
    return *best;
}

void host_stage(rk_server* s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->cv.wait(lk, [&] { return s->stop || !s->waiting.empty(); });
        if (s->stop) return;
        Request r = std::move(s->waiting.front());
        s->waiting.pop_front();
        if (!s->waiting.empty()) s->cv.notify_one();
        if (!r.dev) {
            lk.unlock();
            r.ds->index->filter(r.q, r.k, r.cand);
            lk.lock();
        }
        push_ready(s, std::move(r));
    }
}

std::shared_ptr<rk::PinnedIds> pooled_area(const std::shared_ptr<AreaPool>& pool, size_t n);

// A step slot (0, 1) or the re-run object (2), made on first use with its
// fixed layout, then pointed at the dataset of each step.
rk_batch* object_for(rk_server* s, int which, rk_dataset* ds, rk_status* st)
{
    std::unique_ptr<rk_batch>& b = which < 2 ? s->slots[which] : s->rerun;
    *st = RK_OK;
    if (!b) {
        std::unique_ptr<rk_batch> n(new rk_batch());
        n->ds = ds;
        n->direct = true;
        const std::shared_ptr<AreaPool> pool = s->areas;
        n->area_source = [pool](size_t k) { return pooled_area(pool, k); };
        const ServeLayout& L = s->layout;
        *st = which < 2 ? rk::batch_fix(n.get(), L.nq, L.seg, L.cand) : rk::batch_fix(n.get(), 1, L.re_seg, L.re_cand);
        if (*st != RK_OK) return nullptr;
        b = std::move(n);
    }
    b->ds = ds;
    return b.get();
}

// A step from its enqueue to its publication.
struct Step {
    std::vector<Request> req; // kept until published: re-runs use their candidate lists
    std::vector<rk_query_desc> qd;
    rk_dataset* ds = nullptr;
    bool dev = false;
    rk_batch* b = nullptr;
    bool enqueued = false;
    rk_status st = RK_OK;
    double t_start = 0, t_area = 0, t_filled = 0, t_enqueued = 0; // phase clocks for RK_SERVE_TRACE
};

// RK_SERVE_TRACE=1: a line on stderr for every step with a phase longer than
// a millisecond (fill, enqueue, wait for the run, collect, publish).
bool serve_trace()
{
    static const bool on = [] {
        const char* e = std::getenv("RK_SERVE_TRACE");
        return e && e[0] == '1';
    }();
    return on;
}

// Fills the step's batch object and enqueues its run on the batch's stream.
// The batch objects' device buffers only grow: refilled every step and
// launched without a graph, they allocate nothing once warm.  Host-filtered
// steps time the traversal shapes for a new workload class first;
// device-filtered steps take the default shape (8 waves per query), since
// timing shapes needs filtered lists, which only the step itself makes.
// A landing area of at least n ids for a batch whose own is still held by
// served results: one from the pool, else a new one; it returns to the
// pool when its last reference goes.
std::shared_ptr<rk::PinnedIds> pooled_area(const std::shared_ptr<AreaPool>& pool, size_t n)
{
    rk::PinnedIds* a = nullptr;
    size_t want = n;
    std::vector<rk::PinnedIds*> drop; // free areas below the largest size: freed outside the lock
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        for (size_t i = 0; i < pool->free.size(); ++i)
            if (pool->free[i]->n >= n) {
                a = pool->free[i];
                pool->free.erase(pool->free.begin() + (ptrdiff_t)i);
                break;
            }
        // areas of one size: any free one fits any batch; a new largest
        // size is rounded up to a power of two, so that growing steps
        // allocate on the device stage a few times, not at every new maximum
        want = std::max(n, pool->largest);
        if (!a && want > pool->largest) {
            size_t p2 = 1;
            while (p2 < want) p2 <<= 1;
            want = p2;
        }
        if (!a) pool->largest = want;
        // an area smaller than the largest size fits no batch that holds a
        // larger one, yet it would count as a spare and keep the keeper idle
        for (size_t i = 0; i < pool->free.size();) {
            if (pool->free[i]->n < pool->largest) {
                drop.push_back(pool->free[i]);
                pool->free[i] = pool->free.back();
                pool->free.pop_back();
            } else {
                ++i;
            }
        }
        if (pool->free.size() < kSpare) pool->cv.notify_one();
    }
    for (auto* x : drop) delete x;
    if (!a) {
        a = new rk::PinnedIds();
        if (a->alloc(want) != hipSuccess) {
            delete a;
            return nullptr;
        }
    }
    return std::shared_ptr<rk::PinnedIds>(a, [pool](rk::PinnedIds* x) {
        {
            std::lock_guard<std::mutex> lk(pool->mu);
            if (pool->free.size() < kAreasKeep && x->n >= pool->largest) {
                pool->free.push_back(x);
                return;
            }
        }
        delete x; // freeing pinned memory can wait for the device: not under the pool's lock
    });
}

// The batch gets a landing area no served result refers to.
void fresh_area(rk_server* s, rk_batch* b)
{
    if (!b->land_host || b->land_host.use_count() == 1) return;
    std::shared_ptr<rk::PinnedIds> a = pooled_area(s->areas, b->land_host->n);
    if (a) b->land_host = std::move(a); // else the fill allocates one
    else b->land_host.reset();
}

void start_step(rk_server* s, Step& p, int slot)
{
    p.t_start = rk::now_s();
    p.b = object_for(s, slot, p.ds, &p.st);
    if (!p.b) {
        p.t_area = p.t_filled = p.t_enqueued = rk::now_s();
        return;
    }
    fresh_area(s, p.b);
    p.t_area = rk::now_s();
    if (p.dev) {
        std::vector<uint32_t> q(p.qd.size()), k(p.qd.size());
        for (size_t i = 0; i < p.qd.size(); ++i) {
            q[i] = p.qd[i].q;
            k[i] = p.qd[i].k;
        }
        p.st = rk::batch_fill_device(p.b, q.data(), k.data(), q.size());
    } else {
        p.st = rk::batch_fill(p.b, p.qd.data(), p.qd.size());
        if (p.st == RK_OK) {
            std::lock_guard<std::recursive_mutex> lk(p.ds->ctx->mu);
            p.st = rk::batch_tune(p.b);
        }
    }
    p.t_filled = rk::now_s();
    if (p.st == RK_OK) p.st = rk_batch_enqueue(p.b);
    p.enqueued = p.st == RK_OK;
    p.t_enqueued = rk::now_s();
}

// Waits for the step's run and publishes its results: by reference into the
// batch's landing area when every query fit its segments, else merged with
// the re-runs of the queries that outgrew one (batch_collect) into an array
// of the step's own.
void finish_step(rk_server* s, Step& p)
{
    const size_t nr = p.req.size();
    std::vector<Done> fin(nr);
    rk_status st = p.st;
    const double t_finish = rk::now_s();
    double t_waited = t_finish;
    uint64_t n_ids = 0;
    if (p.enqueued) {
        st = rk_batch_wait(p.b, nullptr);
        t_waited = rk::now_s();
        std::vector<rk::IdRange> ranges;
        std::vector<uint32_t> redo;
        if (st == RK_OK || st == RK_ERR_CAPACITY) {
            std::lock_guard<std::recursive_mutex> lk(p.ds->ctx->mu);
            st = rk::result_ranges(p.b, ranges, &redo);
        }
        if (st == RK_OK && redo.empty()) {
            const std::shared_ptr<const void> area = p.b->land_host;
            for (size_t i = 0; i < nr; ++i) {
                fin[i].ids = ranges[i].src;
                fin[i].n = ranges[i].n;
                // a landed range lives in the area; every-user ranges in the
                // dataset, which outlives the server's use of it
                if (ranges[i].src && p.b->land_host && ranges[i].src >= p.b->land_host->p &&
                    ranges[i].src < p.b->land_host->p + p.b->land_host->n)
                    fin[i].keep = area;
                n_ids += ranges[i].n;
            }
        } else if (st == RK_OK) {
            std::vector<int32_t>* ids_buf = nullptr;
            {
                std::lock_guard<std::mutex> pl(s->pool->mu);
                if (!s->pool->free.empty()) {
                    ids_buf = s->pool->free.back();
                    s->pool->free.pop_back();
                }
            }
            if (!ids_buf) ids_buf = new std::vector<int32_t>();
            std::vector<uint64_t> off(nr + 1);
            rk_batch* r = object_for(s, 2, p.ds, &st);
            if (r) st = rk::batch_collect(p.b, r, p.qd.data(), nr, *ids_buf, off);
            const std::shared_ptr<IdPool> pool = s->pool;
            const std::shared_ptr<const std::vector<int32_t>> shared(ids_buf, [pool](const std::vector<int32_t>* v) {
                auto* w = const_cast<std::vector<int32_t>*>(v);
                std::lock_guard<std::mutex> pl(pool->mu);
                if (pool->free.size() < kPoolKeep) pool->free.push_back(w);
                else delete w;
            });
            if (st == RK_OK)
                for (size_t i = 0; i < nr; ++i) {
                    fin[i].keep = shared;
                    fin[i].ids = shared->data() + off[i];
                    fin[i].n = off[i + 1] - off[i];
                }
            n_ids = off[nr];
        }
    }
    const double t_end = rk::now_s();
    if (serve_trace()) {
        const double ms[4] = {1e3 * (p.t_filled - p.t_start), 1e3 * (p.t_enqueued - p.t_filled),
                              1e3 * (t_waited - t_finish), 1e3 * (t_end - t_waited)};
        if (ms[0] > 1.0 || ms[1] > 1.0 || ms[2] > 1.0 || ms[3] > 1.0)
            std::fprintf(stderr,
                         "rk_serve: slow step: %zu requests, %zu ids; fill %.3f ms (area %.3f), enqueue %.3f ms, "
                         "wait %.3f ms, collect %.3f ms; %.3f ms from enqueue to finish\n",
                         nr, (size_t)n_ids, ms[0], 1e3 * (p.t_area - p.t_start), ms[1], ms[2], ms[3],
                         1e3 * (t_finish - p.t_enqueued));
    }
    for (size_t i = 0; i < nr; ++i) {
        fin[i].status = st;
        fin[i].latency = t_end - p.req[i].t_submit;
        if (st != RK_OK) {
            fin[i].keep.reset();
            fin[i].ids = nullptr;
            fin[i].n = 0;
        }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    for (size_t i = 0; i < nr; ++i) s->done[p.req[i].ticket] = std::move(fin[i]);
    s->cv_done.notify_all();
}

// A request no step can hold (its host candidate list alone outgrows a
// step's layout) is finished at once with RK_ERR_ADMISSION.  Caller holds
// s->mu.
void refuse(rk_server* s, Request&& r)
{
    Done d;
    d.status = RK_ERR_ADMISSION;
    d.latency = rk::now_s() - r.t_submit;
    s->done[r.ticket] = std::move(d);
}

// The device stage keeps up to two steps alive: it fills and enqueues the
// next step, then finishes the previous one, so the device runs a step
// while the host collects and publishes the one before.  With nothing ready
// to overlap, the step in flight is finished at once.
void dev_stage(rk_server* s)
{
    std::unique_ptr<Step> prev; // enqueued, not yet published
    int slot = 0;
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        if (prev && s->n_ready == 0) {
            lk.unlock();
            finish_step(s, *prev);
            prev.reset();
            lk.lock();
            continue;
        }
        s->cv_dev.wait(lk, [&] { return s->stop || s->n_ready > 0; });
        if (s->stop && s->n_ready == 0) return;
        // with the device idle, let batch-mates gather briefly when the step
        // would be small
        if (!prev && s->n_ready < s->cfg.max_batch && s->cfg.max_wait_s > 0 && !s->stop) {
            const double deadline = oldest_ready(s).front().t_submit + s->cfg.max_wait_s;
            while (!s->stop && s->n_ready < s->cfg.max_batch && !s->waiting.empty() && rk::now_s() < deadline) {
                s->cv_dev.wait_for(lk, std::chrono::microseconds(200));
            }
        }
        // one step: the queue (dataset, filter side) whose head waited longest,
        // as many requests as a step slot's layout holds (queries; for host
        // lists also candidates)
        std::deque<Request>& qu = oldest_ready(s);
        std::unique_ptr<Step> cur(new Step());
        cur->ds = qu.front().ds;
        cur->dev = qu.front().dev;
        size_t take = std::min<size_t>(qu.size(), std::min<uint64_t>(s->cfg.max_batch, s->layout.nq));
        if (!cur->dev) {
            const uint64_t room = s->layout.nq * s->layout.cand;
            uint64_t sum = 0;
            size_t t = 0;
            while (t < take && sum + qu[t].cand.size() <= room) sum += qu[t++].cand.size();
            if (t == 0) { // the head's list alone outgrows a step
                refuse(s, std::move(qu.front()));
                qu.pop_front();
                --s->n_ready;
                continue;
            }
            take = t;
        }
        cur->req.reserve(take);
        for (size_t i = 0; i < take; ++i) {
            cur->req.push_back(std::move(qu.front()));
            qu.pop_front();
        }
        s->n_ready -= take;
        lk.unlock();
        cur->qd.resize(take);
        for (size_t i = 0; i < take; ++i) {
            const Request& r = cur->req[i];
            cur->qd[i] = rk_query_desc{r.q, r.k, cur->dev ? nullptr : r.cand.data(), cur->dev ? 0 : r.cand.size()};
        }
        slot ^= 1; // never the batch object of the step in flight
        start_step(s, *cur, slot);
        if (prev) finish_step(s, *prev);
        prev = std::move(cur);
        lk.lock();
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
    s->layout = serve_layout(s->cfg.memory_budget, s->cfg.max_batch);
    const unsigned nf = rk::filter_threads(1u << 20);
    for (unsigned i = 0; i < nf; ++i) s->host_threads.emplace_back(host_stage, s);
    s->dev_thread = std::thread(dev_stage, s);
    s->keeper = std::thread(area_keeper, s->areas);
    *out = s;
    return RK_OK;
}

rk_status rk_serve_submit(rk_server* s, rk_dataset* ds, uint32_t q, uint32_t k, uint64_t* ticket)
{
    if (!s || !ds || !ticket) return RK_ERR_INVALID_ARG;
    if (ds->ctx != s->ctx) return rk::fail(s->ctx, RK_ERR_INVALID_ARG, "dataset belongs to another context");
    if (q >= ds->nf || k == 0) return rk::fail(s->ctx, RK_ERR_INVALID_ARG, "q must be < N_f and k >= 1");
    if (!s->layout.ok)
        return rk::fail(s->ctx, RK_ERR_ADMISSION, "the serving memory budget holds no step; no request can run");
    Request r;
    r.ds = ds;
    r.q = q;
    r.k = k;
    r.dev = !s->host_filter_only && k <= rk::kDfMaxK;
    r.t_submit = rk::now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return RK_ERR_INVALID_ARG;
    r.ticket = s->next_ticket++;
    *ticket = r.ticket;
    if (r.dev) { // nothing to filter on the host: straight to the device stage
        push_ready(s, std::move(r));
        return RK_OK;
    }
    s->waiting.push_back(std::move(r));
    s->cv.notify_one();
    return RK_OK;
}

rk_status rk_serve_poll(rk_server* s, uint64_t ticket, int32_t* out_ids, size_t cap, size_t* n_out, double* latency_s)
{
    if (!s || !n_out || (cap && !out_ids)) return RK_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->done.find(ticket);
    if (it == s->done.end()) {
        if (ticket == 0 || ticket >= s->next_ticket) return RK_ERR_INVALID_ARG;
        return RK_ERR_NOT_READY;
    }
    Done d = it->second;
    *n_out = d.n;
    if (latency_s) *latency_s = d.latency;
    if (d.status != RK_OK) {
        s->done.erase(it);
        return d.status;
    }
    if (d.n > cap) return RK_ERR_CAPACITY; // keep the result for a larger buffer
    s->done.erase(it);
    lk.unlock(); // the copy runs outside the server's lock
    if (d.n) std::memcpy(out_ids, d.ids, d.n * sizeof(int32_t));
    return RK_OK;
}

rk_status rk_serve_info(rk_server* s, uint64_t* device_bytes, uint32_t* step_queries)
{
    if (!s) return RK_ERR_INVALID_ARG;
    if (step_queries) *step_queries = (uint32_t)std::min<uint64_t>(s->layout.nq, s->cfg.max_batch);
    if (device_bytes) {
        // the batch objects belong to the device stage; read them under the
        // context lock, which their fills and runs take
        std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
        uint64_t b = 0;
        for (const auto& o : s->slots)
            if (o) b += rk::batch_device_bytes(o.get());
        if (s->rerun) b += rk::batch_device_bytes(s->rerun.get());
        *device_bytes = b;
    }
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
    {
        std::lock_guard<std::mutex> lk(s->areas->mu);
        s->areas->stop = true;
    }
    s->areas->cv.notify_all();
    if (s->keeper.joinable()) s->keeper.join();
    delete s;
    return RK_OK;
}

} // extern "C"
