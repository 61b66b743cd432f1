// rk_internal.hpp -- library-private object layouts behind the opaque
// handles of include/rknn.h.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <functional>
#include <algorithm>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "rk_common.hpp"
#include "rk_dfilter.hpp"
#include "rk_filter.hpp"
#include "rk_gang.hpp"
#include "rk_kernels.hpp"
#include "rknn.h"

namespace rk {

inline double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Owning device allocation that only grows.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // contents are not kept; a regrowth takes at least 1.5x (buffers refilled
    // step after step settle after a few allocations, which synchronise)
    hipError_t reserve(size_t want)
    {
        if (want <= n && p) return hipSuccess;
        if (p) want = std::max(want, n + n / 2);
        release();
        if (want == 0) want = 1;
        hipError_t e = hipMalloc((void**)&p, want * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        n = want;
        return hipSuccess;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Owning pinned host allocation that only grows; dev is the address the
// kernels reach it by (pinned memory is mapped into the device's space).
template <class T>
struct HostBuf {
    T* p = nullptr;
    T* dev = nullptr;
    size_t n = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf()
    {
        if (p) (void)hipHostFree(p);
    }
    // as DevBuf::reserve: contents not kept, a regrowth takes at least 1.5x
    hipError_t reserve(size_t want, unsigned flags = hipHostMallocDefault)
    {
        if (want <= n && p) return hipSuccess;
        if (p) want = std::max(want, n + n / 2);
        if (p) (void)hipHostFree(p);
        p = dev = nullptr;
        n = 0;
        if (want == 0) want = 1;
        hipError_t e = hipHostMalloc((void**)&p, want * sizeof(T), flags);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        e = hipHostGetDevicePointer((void**)&dev, p, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(p);
            p = dev = nullptr;
            return e;
        }
        n = want;
        return hipSuccess;
    }
};

// Pinned, coherent host ids that a batch's kernels write (the host side of
// its landing area).  Shared: the serving loop hands a step's
// results out by reference and gives the batch another area for its next
// step.
struct PinnedIds {
    int32_t* p = nullptr;   // host address
    int32_t* dev = nullptr; // the address the kernels write through
    size_t n = 0;
    PinnedIds() = default;
    PinnedIds(const PinnedIds&) = delete;
    PinnedIds& operator=(const PinnedIds&) = delete;
    ~PinnedIds()
    {
        if (p) (void)hipHostFree(p);
    }
    hipError_t alloc(size_t ids)
    {
        hipError_t e = hipHostMalloc((void**)&p, std::max<size_t>(ids, 1) * sizeof(int32_t), hipHostMallocCoherent);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        e = hipHostGetDevicePointer((void**)&dev, p, 0);
        if (e != hipSuccess) return e;
        n = std::max<size_t>(ids, 1);
        return hipSuccess;
    }
};

} // namespace rk

struct rk_ctx {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::recursive_mutex mu; // serialises device work on this context
    std::string err;
    std::string arch;        // gcnArchName without feature suffixes (autotuner classes)
    // device address of the allocation-free form's scratch array on this
    // context's device (hipGetSymbolAddress at rk_init; the array's name in
    // host code is not a device address)
    unsigned char* scratch = nullptr;
    // device-to-host copies into pageable memory: two pinned chunks whose
    // DMA overlaps the host threads' copy of the previous chunk
    rk::HostBuf<unsigned char> ring;
    hipEvent_t ring_ev[2] = {nullptr, nullptr};
    // host threads for large host copies (started on first use; one job at
    // a time, under the context lock)
    std::unique_ptr<rk::Gang> gang;
};

// The context's host threads, started on first use.  Caller holds the
// context lock.
inline rk::Gang& host_gang(rk_ctx* c)
{
    if (!c->gang) c->gang.reset(new rk::Gang(std::min(8u, std::max(1u, std::thread::hardware_concurrency()))));
    return *c->gang;
}

struct rk_dataset {
    rk_ctx* ctx = nullptr;
    uint64_t nu = 0, nf = 0;
    uint32_t n_tiles = 0, n_pad = 0;
    rk::DevBuf<float> ux, uy;        // Hilbert-sorted, padded
    rk::DevBuf<int32_t> uid;
    rk::DevBuf<uint32_t> perm;       // sorted position -> caller's user index
    rk::DevBuf<float4> mbr[rk::kMaxLevels];
    rk::DevBuf<float4> gmbr4, gmbr6, gmbr8; // joint rectangles of 4-, 6- and 8-tile groups
    rk::DevBuf<float> fx, fy;        // facilities, caller order
    std::vector<float> hfx, hfy;
    std::vector<int32_t> hfid;
    std::vector<int32_t> all_ids_sorted; // result of a query where every user qualifies
    double ubox[4] = {0, 0, -1, -1};     // users' bounding box (minx, miny, maxx, maxy)
    uint32_t id_bits = 32;               // ids are non-negative and below 2^id_bits (32: any int32)
    std::unique_ptr<rk::FacilityIndex> index;
    rk::DsDev view{};
    // facility grid of the device filter
    rk::FacGridDev fgrid{};
    rk::DevBuf<uint32_t> g_start, g_fid;
    rk::DevBuf<float> g_fx, g_fy;
    rk::DevBuf<double> g_dir;
    // scratch of the plain / count-all forms
    rk::DevBuf<rk::CandRec> rec;
    rk::DevBuf<float2> orig;
    rk::DevBuf<int32_t> out;
    rk::DevBuf<uint32_t> ctr;
    rk::DevBuf<uint32_t> counts, counts_orig;
    rk::DevBuf<unsigned char> counts_many; // rk_count_all_many: one round's count slices
    rk::DevBuf<rk::CountSetQ> setq;        // rk_count_all_many: one round's query-set descriptors
    rk::HostBuf<rk::CountSetQ> h_setq;
    rk::HostBuf<rk::CandRec> h_rec;
    rk::HostBuf<float2> h_orig;
    rk::HostBuf<int32_t> h_out;
    rk::HostBuf<uint32_t> h_ctr;
    // rk_query_batch's batch objects ([0] device-filtered, [1] host-filtered),
    // refilled call after call so that their buffers are allocated once; a
    // call that finds one busy (another thread) uses a batch of its own
    std::mutex qb_mu[2];
    std::unique_ptr<rk_batch> qb[2];
    ~rk_dataset();
};

struct rk_batch {
    rk_dataset* ds = nullptr;
    uint32_t nq = 0;
    std::vector<rk::QDev> hq;
    uint64_t total_cand = 0;
    rk::DevBuf<rk::QDev> q;
    rk::DevBuf<uint32_t> cand;
    rk::DevBuf<double2> wb;
    rk::DevBuf<rk::CandRec> rec;
    rk::DevBuf<float2> orig;
    rk::DevBuf<int32_t> seg;     // per-query result segments (spill of results beyond the kernel's LDS)
    rk::DevBuf<uint32_t> qres;   // [0, nq): results per query, [nq, 2 nq): landing offsets
    rk::DevBuf<uint32_t> qstat;
    // landing area: the traversal kernel writes each query's sorted ids to
    // land_host (pinned, coherent host memory) at the offset qres[nq + i]
    // (a radix-sorted result passes through the same range of dland, in
    // HBM, first) and its kQresWords words to h_qres (pinned, coherent);
    // land_ctr: the run's reservation counters
    rk::HostBuf<uint32_t> h_qres;
    std::shared_ptr<rk::PinnedIds> land_host;
    // where a new host side comes from (at least the ids asked for); empty:
    // a plain allocation
    std::function<std::shared_ptr<rk::PinnedIds>(size_t)> area_source;
    rk::DevBuf<int32_t> dland;
    rk::DevBuf<unsigned long long> land_ctr;
    uint64_t land_cap = 0;
    // inputs staged in pinned, coherent host memory by every fill: a batch
    // refilled step after step (direct) has its kernels read them in place,
    // with no copy; a prepared batch (resident) copies them to HBM once
    bool resident = false;
    rk::HostBuf<rk::QDev> s_q;
    rk::HostBuf<double2> s_wb;
    rk::HostBuf<uint32_t> s_cand, s_dq, s_dk;
    // a fixed layout (the serving loop's objects, sized from its memory
    // budget by batch_fix): at most cap_nq queries per fill, result segments
    // of cap_seg ids, cap_cand candidate slots per query (device lists, and
    // the records of host lists: a fill's lists hold cap_nq * cap_cand
    // candidates at most), a landing area of cap_nq * cap_seg ids; device
    // buffers allocated once, never grown
    uint64_t cap_nq = 0;
    uint64_t cap_seg = 0;
    uint32_t cap_cand = 0;
    rk::BatchDev view{};
    uint32_t* h_qcount() const { return h_qres.p; }
    uint32_t* h_qland() const { return h_qres.p + nq; }
    uint32_t* h_qncand() const { return h_qres.p + 2 * (size_t)nq; }
    uint32_t* h_qflags() const { return h_qres.p + 3 * (size_t)nq; }
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool ran = false;
    bool direct = false;         // run launches the kernels without a graph (refilled batches)
    int shape = 0;               // traversal block shape (rk::kShape*), set by the autotuner
    bool overflow = false;       // the last run had a query beyond its result segment
    uint64_t seg_min = 0;        // result segment per query at least this (capped by N_u), set before a fill
    double filter_s = 0.0;
    // device-filtered batch: (q, k) on the device, the filter inside the step
    bool dfilter = false;
    uint32_t dlists = 0;  // device filter: list sizes of the queries' k (dfilter_lists)
    uint32_t cand_cap = 0;       // candidate slots per query (the filter's segment)
    bool cand_overflow = false;  // the last run had a candidate list beyond cand_cap
    rk::DevBuf<uint32_t> dq, dk, dn;
    std::vector<uint32_t> h_dn;
    // rk_batch_enqueue / rk_batch_wait: a run in flight and its events
    // (start and end of the device work, and done: after the count copy)
    bool pending = false;
    bool pending_empty = false; // nothing was launched (no users or no queries)
    hipEvent_t ev_start = nullptr, ev_end = nullptr, ev_done = nullptr;
    // the batch's own stream for its runs (created with the events), so that
    // runs of different batches overlap on the device: the next batch's query
    // blocks take the SIMDs that the previous batch's finished blocks free
    hipStream_t stream = nullptr;
    ~rk_batch()
    {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : {ev_start, ev_end, ev_done})
            if (e) (void)hipEventDestroy(e);
    }
};

inline rk_dataset::~rk_dataset() = default; // rk_batch is complete here

namespace rk {
rk_status fail(rk_ctx* c, rk_status s, const std::string& msg);
rk_status fail_hip(rk_ctx* c, hipError_t e, const char* what, int line);
// candidate list cleaning shared by every form: validates indices, drops q
// and facilities at q's exact position (never strictly closer)
rk_status clean_candidates(rk_dataset* ds, uint32_t q, const uint32_t* cand, size_t nc, std::vector<uint32_t>& out);
// (re)fills a batch with queries; device buffers only grow, so a batch
// refilled step after step allocates nothing once warm
rk_status batch_fill(rk_batch* b, const rk_query_desc* qs, size_t nq);
// (re)fills a batch with (q, k) pairs for the device filter (k <= kDfMaxK);
// the per-query views are built on the device by each run
rk_status batch_fill_device(rk_batch* b, const uint32_t* q, const uint32_t* k, size_t nq);
// after a run: the results of every query whose result (and device
// candidate list) fit its segment, in query order (ids, off[nq + 1]); the
// others get empty ranges and their indices in redo (a result is complete
// or absent, never truncated)
rk_status batch_fetch_partial(rk_batch* b, std::vector<int32_t>& ids, std::vector<uint64_t>& off,
                              std::vector<uint32_t>& redo);
// the results of a batch that ran over qd[0 .. nq) (rk_batch_run returned
// RK_OK or RK_ERR_CAPACITY), in query order: queries that fit their segments
// come from b; the others are run again in redo (a direct batch, refilled)
// and merged in: a result beyond its segment with segments that hold the
// exact counts b measured, a device candidate list beyond its segment
// through the host filter
rk_status batch_collect(rk_batch* b, rk_batch* redo, const rk_query_desc* qd, size_t nq, std::vector<int32_t>& ids,
                        std::vector<uint64_t>& off);
// the device work of one step of a batch (filter stages when device-filtered)
void batch_launch(rk_batch* b, hipStream_t st);
// device bytes of a fixed layout (nq queries, seg ids per segment, cand
// candidate slots per query), and the layout given to a batch: every device
// buffer allocated at once, exactly
uint64_t batch_fixed_bytes(uint64_t nq, uint64_t seg, uint64_t cand);
// result ids a batch's segments hold in all (2^26, or RK_BATCH_POOL_IDS)
uint64_t batch_pool_ids();
rk_status batch_fix(rk_batch* b, uint64_t nq, uint64_t seg, uint32_t cand);
// device bytes a batch holds
uint64_t batch_device_bytes(const rk_batch* b);
// a range of result ids: n ids at src (to be copied to dst)
struct IdRange {
    int32_t* dst;
    const int32_t* src;
    uint64_t n;
};
// after a run: where each query's sorted ids are (its landing-area range, or
// the dataset's list of every user); with skip, the queries that outgrew a
// segment get empty ranges and are listed there.  Caller holds the context
// lock.
rk_status result_ranges(rk_batch* b, std::vector<IdRange>& ranges, std::vector<uint32_t>* skip);
// picks the traversal block shape for a filled batch (autotuner); caller
// holds the context lock
rk_status batch_tune(rk_batch* b);
// uploads the dataset's facility grid for the device filter (rk_dfilter.hip)
rk_status build_facility_grid(rk_dataset* ds);
// autotuner memory (rk_tune.cpp): the variant chosen for a workload class,
// -1 if none yet
std::string device_arch(rk_ctx* c);
int tuned_choice(const std::string& key);
void tuned_store(const std::string& key, int choice, const char* label);
} // namespace rk

#define RK_HIP(ctx, call)                                                                                      \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) return rk::fail_hip((ctx), e_, #call, __LINE__);                                \
    } while (0)
