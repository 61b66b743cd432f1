// rk_filter.hpp -- host influence-zone filter over a packed facility R-tree.
//
// Replaces the bisector-generation loop of rknn-inf/inf.cpp:201-585 (copied
// as rknn-rt/rt.cpp:289-664) and the facility R*-tree it walks
// (rknn-inf/RStarTree2D.cpp:106-201).  The output is a candidate set C that
// is provably sufficient: for every user u,
//   #{f in C : oracle says f closer than q} < k
//       <=>  #{f in F\{q} : oracle says f closer than q} < k,
// including the oracle's binary64 rounding, so the GPU count over C gives
// exactly the brute-force result.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace rk {

struct FilterStats {
    uint64_t visited = 0;   // facilities popped from the nearest-first stream
    uint64_t candidates = 0;
    double radius = 0.0;    // final influence-zone radius bound
};

// Static packed R-tree (sort-tile-recursive bulk load, fanout 16) with an
// incremental nearest-first facility stream.  Counterpart of the insert-built
// RStarTree (rknn-inf/RStarTree2D.h:87-116) used by the reference filter.
class FacilityIndex {
public:
    FacilityIndex(const float* fx, const float* fy, size_t nf, const double user_box[4]);

    // Candidate facility indices for (q, k), nearest to q first.  Facilities
    // at exactly q's position are never candidates (never strictly closer).
    void filter(uint32_t q, uint32_t k, std::vector<uint32_t>& out, FilterStats* st = nullptr) const;

    size_t size() const { return fx_.size(); }
    static constexpr int kSectors = 64;

private:
    struct Node {
        double minx, miny, maxx, maxy;
        uint32_t first, count; // children: nodes of the level below, or facility slots
    };
    std::vector<float> fx_, fy_;
    std::vector<uint32_t> slot_;              // leaf slot -> facility index
    std::vector<std::vector<Node>> levels_;   // levels_[0] = leaves ... back() = root level
    double box_[4];                           // users' bounding box
    double ex_[kSectors + 1], ey_[kSectors + 1]; // sector boundary directions (ex_[S] = ex_[0])
};

// Filters queries (q[i], k[i]) into out[i] on up to `threads` host threads
// (0: the machine's cores, at most 16, or RK_FILTER_THREADS).  Each query is
// filtered exactly as by FacilityIndex::filter.
void filter_many(const FacilityIndex& idx, const uint32_t* q, const uint32_t* k, size_t n,
                 std::vector<std::vector<uint32_t>>& out, unsigned threads = 0);

// Threads filter_many uses for n queries when asked for `threads` (0: default).
unsigned filter_threads(size_t n, unsigned threads = 0);

} // namespace rk
