/*
 * rknn.h -- C ABI of the MI355X (gfx950) RkNN dominance-count engine.
 *
 * Plain C, POD arguments, caller-owned host buffers, status codes.  No
 * function exits the process or throws across this boundary, and there is no
 * CPU fallback: a missing device is RK_ERR_NO_DEVICE.
 *
 * Semantics (every query form returns the same set):
 *
 *   R(q,k) = { u : #{ f in C : D(u,f) < D(u,q) } < k }     (ties won by q)
 *
 * with D the squared Euclidean distance evaluated in binary64 from the fp32
 * coordinates exactly as dx*dx + dy*dy, dx = (double)u.x - (double)p.x.  C is
 * the candidate facility set: all facilities except q (brute-force form),
 * the host influence-zone filter's output, or a caller-supplied index list.
 * Result ids are returned sorted ascending.  Zero candidates return every
 * user (the reference's rknn-inf-gpu/inf_kernel.cu:76-78 returns none).
 *
 * q is a facility INDEX (position in the facility arrays), as in
 * rknn-inf/inf.cpp:218 (fac[q]); k >= 1.
 */
#ifndef RKNN_H
#define RKNN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rk_status {
    RK_OK = 0,
    RK_ERR_NO_DEVICE = 1,     /* no HIP device visible / device index out of range      */
    RK_ERR_INVALID_ARG = 2,   /* bad pointer, q >= N_f, k == 0, bad candidate index ...  */
    RK_ERR_OUT_OF_MEMORY = 3, /* device allocation failed                                */
    RK_ERR_CAPACITY = 4,      /* caller buffer too small; the size needed is reported    */
    RK_ERR_HIP = 5,           /* a HIP runtime call failed (see rk_last_error)          */
    RK_ERR_IO = 6,            /* file missing or malformed                               */
    RK_ERR_ADMISSION = 7,     /* serving: request exceeds the memory budget              */
    RK_ERR_NOT_READY = 8,     /* serving: ticket still in flight                         */
    RK_ERR_INTERNAL = 9
} rk_status;

typedef struct rk_ctx rk_ctx;         /* one device, one HIP stream                  */
typedef struct rk_dataset rk_dataset; /* device-resident users + facilities + index */
typedef struct rk_batch rk_batch;     /* a prepared, graph-captured batch of queries */
typedef struct rk_server rk_server;   /* a continuous-batching request loop          */

/* Per-call phase times in seconds (wall clock on the host, the reference's
 * get_wall_time phases of rknn-inf-gpu/inf.cpp:610-653) plus the device time
 * of the counting kernels from HIP events. */
typedef struct rk_timing {
    double filter_s;     /* host influence-zone filter                       */
    double upload_s;     /* candidate records host -> device                 */
    double kernel_s;     /* counting kernels (HIP events)                    */
    double download_s;   /* result ids device -> host, including the sort   */
    double total_s;      /* whole call                                       */
    uint64_t n_candidates;
    uint64_t n_fp64_rechecks; /* (user, candidate) tests whose FP32 value fell
                                 in the uncertainty band, settled in binary64 */
    uint64_t n_launches;      /* counting kernels launched (count-all calls; a
                                 launch may count several queries)          */
} rk_timing;

/* ---- lifecycle: replaces cuda_init / cuda_cleanup / is_cuda_available
 *      (rknn-inf-gpu/inf_kernel.h:29-36, inf_kernel.cu:129-160) ---- */
rk_status rk_init(int device, rk_ctx** out);
rk_status rk_shutdown(rk_ctx* ctx);
int rk_device_count(void);                 /* 0 when no device (never fails)   */
const char* rk_status_string(rk_status s);
const char* rk_last_error(const rk_ctx* ctx); /* detail of the last failure    */

/* ---- dataset: replaces the per-call H2D copies of users and lines in
 *      gpu_get_rknn_candidates (inf_kernel.cu:88-95) and gpuReverseKNN
 *      (rknn-naive-gpu/rknn_kernel.cu:172-195); uploaded and Hilbert-sorted
 *      once, then device-resident.  The host keeps the facility index used by
 *      the filter (the facility R*-tree of rknn-inf/inf.cpp:174-177). ---- */
rk_status rk_dataset_create(rk_ctx* ctx,
                            const float* ux, const float* uy, const int32_t* uid, size_t nu,
                            const float* fx, const float* fy, const int32_t* fid, size_t nf,
                            rk_dataset** out);
rk_status rk_dataset_destroy(rk_dataset* ds);
/* users, facilities, device bytes held */
rk_status rk_dataset_info(const rk_dataset* ds, uint64_t* nu, uint64_t* nf, uint64_t* device_bytes);

/* ---- host influence-zone filter: replaces the bisector-generation loop of
 *      rknn-inf/inf.cpp:201-585 (= rknn-rt/rt.cpp:289-664).  Writes candidate
 *      facility indices, nearest first.  RK_ERR_CAPACITY if cap < needed
 *      (*nc then holds the count needed). ---- */
rk_status rk_filter(rk_dataset* ds, uint32_t q, uint32_t k,
                    uint32_t* cand, size_t cap, size_t* nc);

/* ---- the same influence-zone filter on the device, one wave per query over
 *      a facility grid kept in HBM (no reference counterpart: the reference
 *      filters on the host only).  For nq queries (q[i], k[i]), k[i] <= 64,
 *      writes query i's candidate facility indices to
 *      cand[i * cap_per_query ...] and their number to nc[i].  The list may
 *      differ from rk_filter's (another visiting order) but is sufficient in
 *      the same sense: every query form gives the same result over either.
 *      RK_ERR_CAPACITY if a list exceeds cap_per_query (nc holds the sizes). ---- */
rk_status rk_filter_device(rk_dataset* ds, const uint32_t* q, const uint32_t* k, size_t nq,
                           uint32_t* cand, size_t cap_per_query, uint64_t* nc);

/* Same filter on host arrays, no device involved (builds the facility index
 * per call).  Users only bound the zone (their bounding box). */
rk_status rk_filter_points(const float* fx, const float* fy, size_t nf,
                           const float* ux, const float* uy, size_t nu,
                           uint32_t q, uint32_t k, uint32_t* cand, size_t cap, size_t* nc);

/* ---- plain form, one counting launch per query: replaces
 *      gpu_get_rknn_candidates (rknn-inf-gpu/inf_kernel.h:22-27) and the
 *      OptiX launch of rknn-rt (rt.cpp:1116-1148, Params rt.h:40-49).
 *      cand/nc is the candidate list (facility indices, e.g. rk_filter's
 *      output).  out_ids receives the sorted result ids; RK_ERR_CAPACITY if
 *      cap < |R| (*n_out then holds |R|). ---- */
rk_status rk_query_plain(rk_dataset* ds, uint32_t q, uint32_t k,
                         const uint32_t* cand, size_t nc,
                         int32_t* out_ids, size_t cap, size_t* n_out, rk_timing* t);

/* ---- allocation-free plain form.  No counterpart in the reference: it is
 *      for callers that cannot allocate device memory (code that must not
 *      call hipMalloc: serving loops with a fixed footprint, harnesses that
 *      forbid allocation).  Same semantics and result as rk_query_plain; the
 *      kernels' scratch (candidate indices, their records, counters, the
 *      staging of the result ids) is one fixed array per device, set up by
 *      rk_init and held until the process ends (about 4.3 MB).  No call
 *      allocates device or pinned memory.  Limits: nc <= 4096
 *      (RK_ERR_INVALID_ARG beyond); a result counted on the device holds at
 *      most 1,048,576 ids (RK_ERR_CAPACITY with *n_out = |R| beyond, as when
 *      cap < |R|; a query with fewer live candidates than k returns every
 *      user from the host copy, at any size).  Calls on one device
 *      share the array and the library serialises them (a per-device lock);
 *      each call clears what it needs, so a call never sees an earlier call's
 *      contents. ---- */
#define RK_NOALLOC_MAX_CANDIDATES 4096u
#define RK_NOALLOC_MAX_IDS 1048576u
rk_status rk_query_plain_noalloc(rk_dataset* ds, uint32_t q, uint32_t k,
                                 const uint32_t* cand, size_t nc,
                                 int32_t* out_ids, size_t cap, size_t* n_out);

/* ---- brute-force form: C = all facilities but q; replaces gpuReverseKNN
 *      (rknn-naive-gpu/rknn_kernel.cuh:40-46) without its N_u x N_f matrix. ---- */
rk_status rk_query_bruteforce(rk_dataset* ds, uint32_t q, uint32_t k,
                              int32_t* out_ids, size_t cap, size_t* n_out, rk_timing* t);

/* ---- filter + plain: the whole rknn-inf-gpu pipeline (inf.cpp:201-653). ---- */
rk_status rk_query(rk_dataset* ds, uint32_t q, uint32_t k,
                   int32_t* out_ids, size_t cap, size_t* n_out, rk_timing* t);

/* ---- count-all: per-user count of candidates strictly closer than q, no
 *      early exit (the rknn-rt any-hit counter rt.cu:95-111 with k = inf).
 *      counts[i] belongs to the i-th user of rk_dataset_create's arrays.
 *      counts may be host memory or device memory of the context's device
 *      (as rknn-rt's Params::rslt, rt.h:40-49): device counts never cross
 *      PCIe, which bounds the call otherwise (4 bytes per user).  The same
 *      holds for rk_count_all_sorted and rk_count_all_many. ---- */
rk_status rk_count_all(rk_dataset* ds, uint32_t q, const uint32_t* cand, size_t nc,
                       uint32_t* counts, rk_timing* t);

/* ---- count-all in the dataset's own user order: counts[j] belongs to user
 *      order[j] of rk_dataset_create's arrays, order from rk_dataset_order.
 *      The counts leave the count kernel as they are, with no scatter pass
 *      to caller order (a random 4-byte write per user).  Same counts as
 *      rk_count_all (the rknn-rt counter rt.cu:95-111 with k = inf). ---- */
rk_status rk_count_all_sorted(rk_dataset* ds, uint32_t q, const uint32_t* cand, size_t nc,
                              uint32_t* counts, rk_timing* t);

/* ---- the dataset's user order: order[j] = index in rk_dataset_create's
 *      arrays of the j-th user in the engine's layout (nu entries).  Fixed
 *      for the dataset's life: fetch once, apply to any number of
 *      rk_count_all_sorted results.  No reference counterpart. ---- */
rk_status rk_dataset_order(const rk_dataset* ds, uint32_t* order);

/* ---- batched form: many queries, uneven candidate counts and mixed k in
 *      one graph-captured launch sequence.  No reference counterpart beyond
 *      looping gpu_get_rknn_candidates per query. ---- */
typedef struct rk_query_desc {
    uint32_t q;            /* facility index                                */
    uint32_t k;            /* >= 1                                          */
    const uint32_t* cand;  /* candidate facility indices; NULL = run filter */
    uint64_t nc;
} rk_query_desc;

/* ---- count-all over many queries in one call: consecutive queries whose
 *      lists fit 1,024 records together (at most 16 of them) are counted in
 *      one launch, each wave loading its users once for the whole set; a
 *      longer list gets a launch of its own (RK_COUNT_SET=1: every query
 *      does).  The launches run back to back, no host round trip between
 *      them, in rounds of at most 256 MiB of device counts; then each
 *      query's counts are copied out: counts + i * stride (stride >= nu)
 *      receives query i's nu counts, in caller order, or in the dataset's
 *      order (rk_dataset_order) when sorted_order != 0.  A query with cand
 *      NULL counts over the host filter's list at its k; otherwise k is
 *      not used.  t->kernel_s is the time of all the count launches (HIP
 *      events around each round's launches), t->n_launches their number,
 *      t->n_candidates the sum of the lists.  Same counts as rk_count_all
 *      per query (the rknn-rt counter rt.cu:95-111 with k = inf, looped
 *      over queries as rt.cpp:1121-1139 launches one query at a time). ---- */
rk_status rk_count_all_many(rk_dataset* ds, const rk_query_desc* qs, size_t nq, uint32_t* counts,
                            size_t stride, int sorted_order, rk_timing* t);

/* One call: fill + run + fetch.  offsets has nq+1 entries; the ids of
 * query i are out_ids[offsets[i] .. offsets[i+1]), sorted ascending.  The
 * kernels are launched directly on a batch object the dataset keeps per
 * filter side (allocated once).  A query whose result outgrows its share of
 * the batch's result pool, or whose device candidate list outgrows its
 * segment, is run again on its own (exact-size segments, or the host
 * filter); the others' results are kept.  RK_ERR_CAPACITY only when cap is
 * below offsets[nq]. */
rk_status rk_query_batch(rk_dataset* ds, const rk_query_desc* qs, size_t nq,
                         int32_t* out_ids, size_t cap, uint64_t* offsets, rk_timing* t);

/* Split form for resident repeat runs: prepare filters (when cand is NULL),
 * uploads the candidate lists, picks the traversal's block shape (autotuned
 * per workload class: the first batch of a class times each shape on its
 * own data; RK_AUTOTUNE=0 disables, =8 or =4 forces that many waves per
 * query, RK_TUNE_FILE names the JSON cache) and
 * captures the hipGraph; run replays it.  A run ends with every query's
 * sorted ids in the batch's pinned host memory (the step's kernels write
 * them there); fetch copies them into the caller's buffer and
 * rk_batch_results hands out pointers to them.  run returns RK_ERR_CAPACITY
 * when a query's result outgrows its share of the result pool (2^26 ids per
 * batch, RK_BATCH_POOL_IDS); fewer queries per batch give each a larger
 * share. */
rk_status rk_batch_prepare(rk_dataset* ds, const rk_query_desc* qs, size_t nq, rk_batch** out);
rk_status rk_batch_run(rk_batch* b, double* device_seconds);
rk_status rk_batch_fetch(rk_batch* b, int32_t* out_ids, size_t cap, uint64_t* offsets);
/* rk_batch_run in two halves: enqueue puts one run on the batch's own
 * stream and returns; wait blocks until that run is done, then reports its
 * device time (first to last kernel of the run) and checks its segments as
 * rk_batch_run does (RK_ERR_CAPACITY).  Runs of several batches may be in
 * flight at once and overlap on the device, so the device need not idle
 * while the host waits, and one run's blocks fill the SIMDs that another's
 * finished blocks free.  One batch holds at most one run in flight; fetch
 * and stats need its wait first.  No
 * reference counterpart (the reference's kernel call is synchronous,
 * rknn-inf-gpu/inf_kernel.cu:107). */
rk_status rk_batch_enqueue(rk_batch* b);
rk_status rk_batch_wait(rk_batch* b, double* device_seconds);
/* The last run's results in place, without a copy: ids[i] points at query
 * i's n[i] sorted ids in the batch's pinned landing area (the step's
 * kernels wrote them there; or in the dataset's list of every user), valid
 * until the batch is filled, enqueued or destroyed.  Arrays of nq entries;
 * the conditions of rk_batch_fetch apply.  No reference counterpart (the
 * reference copies into the caller's vector, inf_kernel.cu:115-120). */
rk_status rk_batch_results(rk_batch* b, const int32_t** ids, uint64_t* n);
rk_status rk_batch_destroy(rk_batch* b);
/* Device-filtered batch: the batch holds (q[i], k[i]) pairs, k[i] <= 64, and
 * each run filters on the device (rk_filter_device's filter) before the
 * traversal, all in one graph, so a run goes from queries to result ids
 * without the host.  run returns RK_ERR_CAPACITY if a candidate list
 * exceeds its device segment (4096 facilities); rk_query_batch then runs
 * that query again through the host filter.  rk_query_batch uses this form
 * for queries without candidate lists and k <= 64 unless RK_FILTER=host.  No
 * reference counterpart. */
rk_status rk_batch_prepare_device(rk_dataset* ds, const uint32_t* q, const uint32_t* k, size_t nq,
                                  rk_batch** out);
/* Work counters of the last run: items[l] = tree nodes of level l that
 * entered a query's frontier (0 = leaf tiles tested; n_levels entries, at
 * most 8), result ids, (user, candidate) tests settled in binary64. */
rk_status rk_batch_stats(const rk_batch* b, uint64_t* items, size_t n_items_cap, size_t* n_levels,
                         uint64_t* results, uint64_t* rechecks);
/* The block shape the autotuner chose: waves per query (8 or 4).  No
 * reference counterpart. */
rk_status rk_batch_config(const rk_batch* b, uint32_t* waves_per_query);

/* ---- serving: continuous batching over resident datasets. ---- */
typedef struct rk_serve_config {
    uint32_t max_batch;          /* queries per device step (default 256)          */
    uint64_t memory_budget;      /* device bytes the server may allocate (default  *
                                  * 4 GiB): its step and re-run buffers, laid out  *
                                  * once from the budget; steps take only what     *
                                  * that layout holds.  Resident datasets and      *
                                  * pinned host memory are not counted.            */
    double max_wait_s;           /* longest a request waits for batch-mates while   *
                                  * no step is in flight                            */
} rk_serve_config;
/* The server's device thread keeps up to two steps alive: it fills and
 * enqueues the next step, then finishes (waits for, collects, publishes) the
 * previous one, so the device runs a step while the host finishes the one
 * before.  A step in flight is finished at once when nothing else is ready. */

rk_status rk_serve_start(rk_ctx* ctx, const rk_serve_config* cfg, rk_server** out);
/* Thread-safe.  Returns a ticket; RK_ERR_ADMISSION if the budget holds no
 * step at all.  A request is admitted into a step when the step's layout
 * holds it; one that no step could hold (a host candidate list longer than a
 * step's candidate slots, a result beyond the re-run share) completes with
 * RK_ERR_ADMISSION from rk_serve_poll. */
rk_status rk_serve_submit(rk_server* s, rk_dataset* ds, uint32_t q, uint32_t k, uint64_t* ticket);
/* RK_ERR_NOT_READY while in flight; on RK_OK the ticket is retired.
 * latency_s = completion - submission. */
rk_status rk_serve_poll(rk_server* s, uint64_t ticket, int32_t* out_ids, size_t cap,
                        size_t* n_out, double* latency_s);
/* Device bytes the server's buffers hold (at most memory_budget) and the
 * queries a step takes at most.  No reference counterpart. */
rk_status rk_serve_info(rk_server* s, uint64_t* device_bytes, uint32_t* step_queries);
rk_status rk_serve_stop(rk_server* s);

#ifdef __cplusplus
}
#endif
#endif /* RKNN_H */
