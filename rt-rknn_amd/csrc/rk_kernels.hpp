// rk_kernels.hpp -- host-side launchers of the gfx950 kernels (rk_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "rk_common.hpp"

namespace rk {

// Batch-wide device state written by the kernels.
struct BatchCounters {
    uint32_t items[kMaxLevels + 1]; // items pushed per parent level (0 = leaf tiles)
    uint32_t overflow;              // nonzero: an item list or the pool ran out of room
    uint32_t pool;                  // result entries appended
    uint32_t sort_overflow;         // queries whose result did not fit the LDS sort
    uint32_t rechecks;              // users resolved by the binary64 re-check
    uint32_t pad[3];
};

struct BatchDev {
    const QDev* q;          // nq queries
    uint32_t nq;
    const CandRec* rec;     // all candidate records, query-major
    const float2* orig;     // all candidate facility coordinates (for the re-check)
    uint64_t* items[kMaxLevels + 1]; // (query << 32 | node) per parent level
    uint32_t item_cap[kMaxLevels + 1];
    int32_t* pool_uid;      // appended result ids
    uint64_t* pool_tag;     // (query << 32 | rank within query)
    uint32_t pool_cap;
    uint32_t* qcount;       // results per query
    uint64_t* qoff;         // nq+1 exclusive offsets into out
    int32_t* out;           // results grouped per query, then sorted
    BatchCounters* ctr;
};

// dataset construction
void launch_morton_keys(const float* ux, const float* uy, uint32_t n, float minx, float miny, float sx,
                        float sy, uint32_t* keys, uint32_t* idx, hipStream_t st);
void launch_gather_users(const uint32_t* order, uint32_t n, uint32_t n_pad, const float* ux_in,
                         const float* uy_in, const int32_t* uid_in, float* ux, float* uy, int32_t* uid,
                         hipStream_t st);
void launch_tile_mbr(const float* ux, const float* uy, uint32_t n_tiles, float4* mbr, hipStream_t st);
void launch_level_mbr(const float4* child, uint32_t n_child, float4* parent, uint32_t n_parent, hipStream_t st);

// plain form: one launch over every tile for one query
void launch_plain(const DsDev& ds, QDev q, const CandRec* rec, const float2* orig, int32_t* out_ids,
                  uint32_t* counters /* [0]=n_out [1]=rechecks */, int grid, hipStream_t st);

// count-all form: exact per-user count (sorted order), no early exit
void launch_count_all(const DsDev& ds, QDev q, const CandRec* rec, const float2* orig, uint32_t* counts,
                      uint32_t* rechecks, int grid, hipStream_t st);
void launch_scatter_u32(const uint32_t* src_sorted, const uint32_t* perm, uint32_t n, uint32_t* dst_orig,
                        hipStream_t st);

// batched form
void launch_prep_records(const QDev* q, uint32_t nq, const uint32_t* cand, const float* fx, const float* fy,
                         const double2* wbound, CandRec* rec, float2* orig, uint64_t total, hipStream_t st);
void launch_batch_init(const BatchDev& b, uint32_t root_level, hipStream_t st);
void launch_batch_internal(const DsDev& ds, const BatchDev& b, uint32_t parent_level, int grid,
                           hipStream_t st);
void launch_batch_leaf(const DsDev& ds, const BatchDev& b, int grid, hipStream_t st);
void launch_batch_finish(const BatchDev& b, int grid, hipStream_t st);

int device_grid(int n_cu, int waves_per_cu);

} // namespace rk
