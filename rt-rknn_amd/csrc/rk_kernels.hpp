// rk_kernels.hpp -- host-side launchers of the gfx950 kernels (rk_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "rk_common.hpp"

namespace rk {

// Slots of the plain form's per-wave statistics counters: one counter hit by
// every wave would serialise thousands of device atomics.
constexpr int kBuckets = 64;

// Batched form: one block of up to kQueryBlock threads per query; tree frontiers
// of kFront nodes per level in LDS.  A query's result ids stay in LDS (up to
// kResLds of them), are sorted there and written once to the batch's landing
// area in host memory; larger results spill to the query's segment in HBM,
// and a result the block's in-LDS sort does not take is radix-sorted by the
// same block through the segment and the device landing area.
constexpr int kQueryBlock = 512;
constexpr uint32_t kFront = 512;
constexpr uint32_t kResLds = 2048;
// per-query statistics: [0] binary64 tests, [1 + l] nodes entering level l,
// [kSortSlot] nonzero when the block radix-sorted the result
constexpr uint32_t kSortSlot = 1 + kMaxLevels;
constexpr uint32_t kQueryStats = 2 + kMaxLevels;
// landing offset of a query whose ids did not reach the landing area (the
// area was full, or the result outgrew its segment)
constexpr uint32_t kNotLanded = 0xffffffffu;
// the landing area holds fewer ids than kNotLanded
constexpr uint64_t kMaxLandIds = 0xfffffff0ull;
// words per query of BatchDev::hqres, and its flag bits
constexpr uint32_t kQresWords = 4;
constexpr uint32_t kQflagTrivial = 1u; // k > |C|: every user qualifies (no traversal ran)

struct BatchDev {
    const QDev* q;          // nq queries
    uint32_t nq;
    uint32_t seg_cap;       // result ids one query's segment holds
    const uint32_t* cand;   // candidate facility indices, query-major (QDev::off)
    const float* fx;        // facility coordinates of the dataset
    const float* fy;
    const double2* wbound;  // per-query |u - q| axis bounds of the record scaling
    CandRec* rec;           // candidate records of queries with > 1024 candidates
    float2* orig;           // their facility coordinates (for the binary64 re-check)
    int32_t* seg;           // nq segments of seg_cap ids: query i's spill at seg[i * seg_cap]
    uint32_t* qcount;       // results per query (may exceed seg_cap: overflow)
    uint32_t* qland;        // per query: offset of its sorted ids in land[], or kNotLanded
    uint32_t* qstat;        // nq x kQueryStats statistics
    int32_t* dland;         // landing area on the device (radix-sorted results), land_cap ids
    int32_t* hland;         // its host twin (pinned), land_cap ids: every landed query's sorted ids
    uint32_t* hqres;        // pinned host words per query, written by the traversal kernel: [0, nq) result
                            // count, [nq, 2 nq) landing offset, [2 nq, 3 nq) candidates (the device
                            // filter's count, beyond its segment when it overflowed), [3 nq, 4 nq) flags
    const uint32_t* dn;     // device filter's candidate counts (device-filtered batches), else null
    uint64_t land_cap;
    unsigned long long* land_ctr; // [0] ids reserved by this run, [1] traversal blocks done
    uint32_t id_bits;       // user ids are non-negative and below 2^id_bits (32: any int32)
};

// dataset construction
void launch_curve_keys(const float* ux, const float* uy, uint32_t n, float minx, float miny, float sx,
                       float sy, uint32_t* keys, uint32_t* idx, hipStream_t st);
void launch_gather_users(const uint32_t* order, uint32_t n, uint32_t n_pad, const float* ux_in,
                         const float* uy_in, const int32_t* uid_in, float* ux, float* uy, int32_t* uid,
                         hipStream_t st);
void launch_tile_mbr(const float* ux, const float* uy, uint32_t n_tiles, float4* mbr, hipStream_t st);
void launch_level_mbr(const float4* child, uint32_t n_child, float4* parent, uint32_t n_parent, hipStream_t st);
// rectangles of groups of fan consecutive tiles (the last group may be short)
void launch_group_mbr(const float4* tiles, uint32_t n_tiles, uint32_t fan, float4* groups, hipStream_t st);

// plain form: one launch over every tile for one query
// counters: [0] = ids written, [1 .. kBuckets] = binary64 tests per wave slot
constexpr int kPlainCounters = 1 + kBuckets;
// out_ids holds out_cap ids; counters[0] counts every kept user, stored or not
void launch_plain(const DsDev& ds, QDev q, const CandRec* rec, const float2* orig, int32_t* out_ids,
                  uint32_t out_cap, uint32_t* counters, int grid, hipStream_t st);
// records of nc candidate indices for query q (dead records for q and for
// facilities at q's position); also zeroes counters[0 .. n_ctr)
void launch_build_records(const uint32_t* cand, uint32_t nc, uint32_t q, float qx, float qy, double wx, double wy,
                          const float* fx, const float* fy, CandRec* rec, float2* orig, uint32_t* counters,
                          uint32_t n_ctr, hipStream_t st);

// count-all form: exact per-user count (sorted order), no early exit.
// rechecks has kBuckets slots.  Counts take count_bytes(nc) bytes each: 1
// below 2^8 candidates, 2 below 2^16, else 4 (always 4 when wide); the
// buffer holds n_pad of them.
// variants: 4, 6 or 8 users per lane (3 instructions per clean test in
// each), one group of tiles per wave
constexpr int kCountValu4 = 0;
constexpr int kCountValu8 = 1;
constexpr int kCountValu6 = 2;
constexpr int kCountVariants = 3;
const char* count_variant_name(int variant);
int count_bytes(uint32_t nc);
// wide: 32-bit counts whatever the candidate count
void launch_count_all(const DsDev& ds, QDev q, const CandRec* rec, const float2* orig, void* counts,
                      uint32_t* rechecks, int variant, hipStream_t st, bool wide = false);
void launch_scatter_counts(const void* src_sorted, int bytes, const uint32_t* perm, uint32_t n,
                           uint32_t* dst_orig, hipStream_t st);

// count-all over a set of queries in one launch: every wave counts its group
// of tiles against each query of the set.  The set's records (n_recs in all,
// query i's at rec + qs[i].roff) are staged in LDS, so n_recs <=
// kCountSetRecs; at most kCountSetMax queries.  Every query's counts take
// bytes (1, 2 or 4) each: query i's n_pad counts at (char*)counts + bytes *
// qs[i].coff.  qs is device memory.
constexpr uint32_t kCountSetMax = 16;
constexpr uint32_t kCountSetRecs = 1024;
struct CountSetQ {
    float qx, qy;
    uint32_t nc;   // records of the query
    uint32_t roff; // first record in the set's rec[] / orig[]
    uint64_t coff; // first count, in counts
    uint64_t pad;
};
void launch_count_set(const DsDev& ds, const CountSetQ* qs, uint32_t nq, uint32_t n_recs, const CandRec* rec,
                      const float2* orig, void* counts, int bytes, uint32_t* rechecks, int variant, hipStream_t st);

// batched form: the whole step is one launch of the traversal kernel, which
// sorts each query's result and writes its ids and words to pinned host
// memory; captured in a graph.  After the step, hqres says where each query's
// sorted ids are in hland; the kernel's last block clears land_ctr for the
// next run.
size_t batch_query_lds_bytes(uint32_t n_levels);
// block shapes of the traversal kernel (waves per query), chosen by the autotuner
constexpr int kShape8Waves = 0;
constexpr int kShape4Waves = 1;
constexpr int kNumShapes = 2;
void launch_batch_query(const DsDev& ds, const BatchDev& b, int shape, hipStream_t st);

int device_grid(int n_cu, int waves_per_cu);

} // namespace rk
