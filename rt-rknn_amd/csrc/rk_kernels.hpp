// rk_kernels.hpp -- host-side launchers of the gfx950 kernels (rk_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "rk_common.hpp"

namespace rk {

// Batch-wide device state written by the kernels.
// Each level's work list is split into kBuckets regions with their own
// append counters, so concurrent appends spread over kBuckets addresses.
constexpr int kBuckets = 64;

struct BatchCounters {
    uint32_t items[kMaxLevels + 1][kBuckets]; // items appended per parent level and bucket
    uint32_t overflow;              // nonzero: an item list or a result segment ran out of room
    uint32_t results;               // total result ids (written by the scan)
    uint32_t sort_overflow;         // queries whose result did not fit the LDS sort
    uint32_t pad;
    // (user, candidate) tests settled in binary64, one slot per (wave % kBuckets):
    // a single counter hit by every wave serialises thousands of device atomics
    uint32_t rechecks[kBuckets];
};

struct BatchDev {
    const QDev* q;          // nq queries
    uint32_t nq;
    const CandRec* rec;     // all candidate records, query-major
    const float2* orig;     // all candidate facility coordinates (for the re-check)
    uint64_t* items[kMaxLevels + 1]; // (query << 32 | node) per parent level, kBuckets regions
    uint32_t item_cap[kMaxLevels + 1]; // capacity of one bucket region
    int32_t* seg;           // nq segments of seg_cap ids: query i's results at seg[i * seg_cap]
    uint32_t seg_cap;
    uint32_t* qcount;       // results per query
    uint64_t* qoff;         // nq+1 exclusive offsets of the compacted layout
    int32_t* out;           // compacted results (written when fetched)
    BatchCounters* ctr;
};

// dataset construction
void launch_curve_keys(const float* ux, const float* uy, uint32_t n, float minx, float miny, float sx,
                       float sy, uint32_t* keys, uint32_t* idx, hipStream_t st);
void launch_gather_users(const uint32_t* order, uint32_t n, uint32_t n_pad, const float* ux_in,
                         const float* uy_in, const int32_t* uid_in, float* ux, float* uy, int32_t* uid,
                         hipStream_t st);
void launch_tile_mbr(const float* ux, const float* uy, uint32_t n_tiles, float4* mbr, hipStream_t st);
void launch_level_mbr(const float4* child, uint32_t n_child, float4* parent, uint32_t n_parent, hipStream_t st);

// plain form: one launch over every tile for one query
// counters: [0] = ids written, [1 .. kBuckets] = binary64 tests per wave slot
constexpr int kPlainCounters = 1 + kBuckets;
void launch_plain(const DsDev& ds, QDev q, const CandRec* rec, const float2* orig, int32_t* out_ids,
                  uint32_t* counters, int grid, hipStream_t st);

// count-all form: exact per-user count (sorted order), no early exit;
// rechecks has kBuckets slots
void launch_count_all(const DsDev& ds, QDev q, const CandRec* rec, const float2* orig, uint32_t* counts,
                      uint32_t* rechecks, int grid, hipStream_t st);
void launch_scatter_u32(const uint32_t* src_sorted, const uint32_t* perm, uint32_t n, uint32_t* dst_orig,
                        hipStream_t st);

// batched form
void launch_batch_prep(const BatchDev& b, const uint32_t* cand, const float* fx, const float* fy,
                       const double2* wbound, uint64_t total, hipStream_t st);
void launch_batch_root(const DsDev& ds, const BatchDev& b, int grid, hipStream_t st);
void launch_batch_internal(const DsDev& ds, const BatchDev& b, uint32_t parent_level, int grid,
                           hipStream_t st);
void launch_batch_leaf(const DsDev& ds, const BatchDev& b, int grid, hipStream_t st);
void launch_batch_sort(const BatchDev& b, hipStream_t st);
void launch_batch_scan(const BatchDev& b, hipStream_t st);    // offsets of the compacted layout
void launch_batch_compact(const BatchDev& b, hipStream_t st); // segments -> compacted layout

int device_grid(int n_cu, int waves_per_cu);

} // namespace rk
