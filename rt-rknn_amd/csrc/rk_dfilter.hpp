// rk_dfilter.hpp -- device influence-zone filter (rk_dfilter.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "rk_common.hpp"

namespace rk {

constexpr int kDfSectors = 64;     // angular sectors around q (one per lane)
constexpr uint32_t kDfMaxK = 64;   // k the device filter takes (sector heaps in LDS)
constexpr uint32_t kDfBatchCand = 4096; // candidate slots per query of a device-filtered batch

// Facility grid of a dataset: a uniform grid over the facilities' bounding
// box, about two facilities per cell, cells in row-major order.
struct FacGridDev {
    const float* fx;        // coordinates sorted by cell
    const float* fy;
    const uint32_t* fid;    // facility index (caller order) of each sorted slot
    const uint32_t* start;  // G*G + 1 offsets into the sorted arrays
    const float* ofx;       // coordinates in caller order (q's position)
    const float* ofy;
    const double* ex;       // kDfSectors + 1 sector boundary directions (last = first)
    const double* ey;
    double x0, y0, cs;      // grid origin and cell side
    uint32_t G, nf;
    double box[4];          // users' bounding box (minx, miny, maxx, maxy)
};

// One wave per query: candidate facility indices of query i at
// cand[i * cap ..], their number in ncand[i] (may exceed cap: overflow).
// Entries of a query's sector lists: 16, 32 or 64, the fewest that hold k.
RK_HD inline uint32_t dfilter_list(uint32_t k) { return k <= 16 ? 16u : k <= 32 ? 32u : 64u; }
// Bit 0, 1, 2: some k of k[0 .. n) takes lists of 16, 32, 64 entries.
uint32_t dfilter_lists(const uint32_t* k, size_t n);
// lists: dfilter_lists of the queries' k (a kernel of that list size when
// there is one, else one for every size)
void launch_dfilter(const FacGridDev& g, const uint32_t* q, const uint32_t* k, uint32_t nq, uint32_t lists,
                    uint32_t* cand, uint32_t cap, uint32_t* ncand, hipStream_t st);
// QDev and record-scaling bounds of a device-filtered batch from the counts
void launch_dfilter_setup(const FacGridDev& g, const uint32_t* q, const uint32_t* k, uint32_t nq,
                          const uint32_t* ncand, uint32_t cap, uint32_t nu, QDev* qd, double2* wb, hipStream_t st);

} // namespace rk
