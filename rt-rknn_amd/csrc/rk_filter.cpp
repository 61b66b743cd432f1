// rk_filter.cpp -- conservative influence-zone filter (see rk_filter.hpp).
//
// The reference filter (rknn-inf/inf.cpp:201-585) keeps the vertices of the
// bisector arrangement and prunes a facility when no "considered" vertex is
// closer to it than to q (inf.cpp:318-353).  This filter bounds the same
// influence zone with angular sectors around q instead, a partition in the
// spirit of rknn-slice (slice.h:306-376):
//
//  * Along a ray q + r*e, facility f's half-plane H_f = {p : f strictly
//    closer than q} is entered at r = rho_f(e) = |g|^2 / (2 e.g), g = f - q,
//    and never left (infinite when e.g <= 0).
//  * In a sector bounded by directions e_a, e_b (< 180 degrees), rho_f is
//    largest at an end ray and smallest at g's direction if it lies inside,
//    else at the nearer end ray.
//  * If k candidates have sector-max rho <= B_s, every point of the sector
//    beyond B_s lies strictly inside k half-planes, so the zone where fewer
//    than k candidates are closer (where every result user lives) is within
//    distance B_s of q in that sector; B_s is also capped by the farthest
//    point of the users' bounding box in the sector.
//  * A facility whose sector-min rho is >= B_s in every sector cannot be
//    closer than q to any user of that zone and is pruned.  Facilities come
//    nearest-first, and the stream stops once |f-q|/2 >= max_s B_s.
//
// Rounding: half-planes used to shrink the bounds are tightened and
// half-planes tested for pruning are loosened by the binary64 oracle's own
// error (4.01*2^-53 relative on each squared distance, doubled), so the
// candidate set is sufficient for the oracle's decisions, not only for exact
// geometry.  Extra candidates are always harmless.
#include "rk_filter.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <numeric>
#include <queue>
#include <thread>

namespace rk {

namespace {

constexpr int kFan = 16;
constexpr double kInf = std::numeric_limits<double>::infinity();

// Sort-tile-recursive order of items with centres (cx, cy), grouped by kFan.
template <class GetX, class GetY>
std::vector<uint32_t> str_order(size_t n, GetX gx, GetY gy)
{
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return gx(a) < gx(b); });
    const size_t groups = (n + kFan - 1) / kFan;
    const size_t slabs = (size_t)std::ceil(std::sqrt((double)groups));
    const size_t slab = slabs ? slabs * kFan : n;
    for (size_t s = 0; s < n; s += slab) {
        auto b = idx.begin() + s, e = idx.begin() + std::min(n, s + slab);
        std::sort(b, e, [&](uint32_t a, uint32_t c) { return gy(a) < gy(c); });
    }
    return idx;
}

inline double mindist2(double qx, double qy, double minx, double miny, double maxx, double maxy)
{
    double dx = qx < minx ? minx - qx : (qx > maxx ? qx - maxx : 0.0);
    double dy = qy < miny ? miny - qy : (qy > maxy ? qy - maxy : 0.0);
    return dx * dx + dy * dy;
}

} // namespace

FacilityIndex::FacilityIndex(const float* fx, const float* fy, size_t nf, const double user_box[4])
    : fx_(fx, fx + nf), fy_(fy, fy + nf)
{
    for (int i = 0; i < 4; ++i) box_[i] = user_box[i];
    const double two_pi = 6.283185307179586476925286766559;
    for (int s = 0; s < kSectors; ++s) {
        ex_[s] = std::cos(two_pi * s / kSectors);
        ey_[s] = std::sin(two_pi * s / kSectors);
    }
    ex_[kSectors] = ex_[0];
    ey_[kSectors] = ey_[0];
    if (nf == 0) return;

    // leaves over facility slots
    std::vector<uint32_t> ord =
        str_order(nf, [&](uint32_t i) { return (double)fx_[i]; }, [&](uint32_t i) { return (double)fy_[i]; });
    slot_ = ord;
    std::vector<Node> level;
    for (size_t s = 0; s < nf; s += kFan) {
        Node n{kInf, kInf, -kInf, -kInf, (uint32_t)s, (uint32_t)std::min<size_t>(kFan, nf - s)};
        for (uint32_t i = n.first; i < n.first + n.count; ++i) {
            const double x = fx_[slot_[i]], y = fy_[slot_[i]];
            n.minx = std::min(n.minx, x);
            n.miny = std::min(n.miny, y);
            n.maxx = std::max(n.maxx, x);
            n.maxy = std::max(n.maxy, y);
        }
        level.push_back(n);
    }
    // upper levels: reorder the level by STR, then group consecutive nodes
    while (true) {
        if (level.size() <= (size_t)kFan) {
            levels_.push_back(level);
            break;
        }
        std::vector<uint32_t> o = str_order(
            level.size(), [&](uint32_t i) { return 0.5 * (level[i].minx + level[i].maxx); },
            [&](uint32_t i) { return 0.5 * (level[i].miny + level[i].maxy); });
        std::vector<Node> sorted(level.size());
        for (size_t i = 0; i < o.size(); ++i) sorted[i] = level[o[i]];
        std::vector<Node> parent;
        for (size_t s = 0; s < sorted.size(); s += kFan) {
            Node n{kInf, kInf, -kInf, -kInf, (uint32_t)s, (uint32_t)std::min<size_t>(kFan, sorted.size() - s)};
            for (uint32_t i = n.first; i < n.first + n.count; ++i) {
                n.minx = std::min(n.minx, sorted[i].minx);
                n.miny = std::min(n.miny, sorted[i].miny);
                n.maxx = std::max(n.maxx, sorted[i].maxx);
                n.maxy = std::max(n.maxy, sorted[i].maxy);
            }
            parent.push_back(n);
        }
        levels_.push_back(std::move(sorted));
        level = std::move(parent);
    }
}

void FacilityIndex::filter(uint32_t q, uint32_t k, std::vector<uint32_t>& out, FilterStats* st) const
{
    out.clear();
    const size_t nf = fx_.size();
    if (nf == 0 || q >= nf || k == 0) return;
    const double qx = fx_[q], qy = fy_[q];
    const double bx0 = box_[0], by0 = box_[1], bx1 = box_[2], by1 = box_[3];
    if (!(bx0 <= bx1 && by0 <= by1)) return; // no users

    // farthest squared distance from q to the users' box (bounds D(u,q))
    double rbox2 = 0.0;
    const double cx[4] = {bx0, bx1, bx0, bx1}, cy[4] = {by0, by0, by1, by1};
    for (int i = 0; i < 4; ++i) {
        const double dx = cx[i] - qx, dy = cy[i] - qy;
        rbox2 = std::max(rbox2, dx * dx + dy * dy);
    }
    // binary64 oracle error on D(u,f) - D(u,q): 4.01*2^-53*(D_f + D_q) with
    // D_f + D_q <= 3 D_q + 2|g|^2; doubled for the evaluation of these bounds
    auto delta = [&](double gg) { return 8.02 * 0x1p-53 * (3.0 * rbox2 + 2.0 * gg) + 0x1p-100; };

    // per-sector cap: farthest point of (users' box) within the sector
    const int S = kSectors;
    double cap[kSectors];
    const double tol = 1e-12;
    for (int s = 0; s < S; ++s) {
        const double ax = ex_[s], ay = ey_[s], bx = ex_[s + 1], by = ey_[s + 1];
        double best = -1.0;
        if (qx >= bx0 && qx <= bx1 && qy >= by0 && qy <= by1) best = 0.0;
        for (int i = 0; i < 4; ++i) {
            const double dx = cx[i] - qx, dy = cy[i] - qy;
            const double sc = tol * (std::fabs(dx) + std::fabs(dy));
            if (ax * dy - ay * dx >= -sc && dx * by - dy * bx >= -sc) best = std::max(best, dx * dx + dy * dy);
        }
        const double rx[2] = {ax, bx}, ry[2] = {ay, by};
        for (int r = 0; r < 2; ++r) {
            const double ux = rx[r], uy = ry[r];
            const double span = (bx1 - bx0) + (by1 - by0) + 1.0;
            if (ux != 0.0) {
                for (double X : {bx0, bx1}) {
                    const double t = (X - qx) / ux;
                    const double y = qy + t * uy;
                    if (t >= 0.0 && y >= by0 - tol * span && y <= by1 + tol * span)
                        best = std::max(best, t * t);
                }
            }
            if (uy != 0.0) {
                for (double Y : {by0, by1}) {
                    const double t = (Y - qy) / uy;
                    const double x = qx + t * ux;
                    if (t >= 0.0 && x >= bx0 - tol * span && x <= bx1 + tol * span)
                        best = std::max(best, t * t);
                }
            }
        }
        cap[s] = best < 0.0 ? 0.0 : std::sqrt(best) * (1.0 + 1e-9) + 1e-9;
    }

    // per-sector max-heaps of the k smallest tightened sector-max rho
    std::vector<std::vector<double>> heap(S);
    double bound[kSectors];
    double R = 0.0;
    for (int s = 0; s < S; ++s) {
        bound[s] = cap[s];
        R = std::max(R, bound[s]);
    }

    // nearest-first stream over the packed tree
    struct Ent {
        double key;
        int32_t level; // -1: facility slot
        uint32_t idx;
        bool operator>(const Ent& o) const { return key > o.key; }
    };
    std::priority_queue<Ent, std::vector<Ent>, std::greater<Ent>> pq;
    const int top = (int)levels_.size() - 1;
    for (uint32_t i = 0; i < levels_[top].size(); ++i) {
        const Node& n = levels_[top][i];
        pq.push({mindist2(qx, qy, n.minx, n.miny, n.maxx, n.maxy), top, i});
    }
    double ca[kSectors + 1], xs[kSectors + 1];
    uint64_t visited = 0;
    while (!pq.empty()) {
        const Ent e = pq.top();
        pq.pop();
        // everything still queued is at least sqrt(e.key) from q
        if (e.key > 0.0) {
            const double dl = delta(e.key) / e.key;
            const double lower = dl >= 1.0 ? 0.0 : 0.5 * std::sqrt(e.key) * (1.0 - dl) * (1.0 - 1e-12);
            if (lower >= R) break;
        }
        if (e.level >= 0) {
            const Node& n = levels_[e.level][e.idx];
            if (e.level == 0) {
                for (uint32_t i = n.first; i < n.first + n.count; ++i) {
                    const uint32_t f = slot_[i];
                    const double dx = (double)fx_[f] - qx, dy = (double)fy_[f] - qy;
                    pq.push({dx * dx + dy * dy, -1, f});
                }
            } else {
                const std::vector<Node>& lv = levels_[e.level - 1];
                for (uint32_t i = n.first; i < n.first + n.count; ++i)
                    pq.push({mindist2(qx, qy, lv[i].minx, lv[i].miny, lv[i].maxx, lv[i].maxy), e.level - 1, i});
            }
            continue;
        }
        const uint32_t f = e.idx;
        ++visited;
        if (f == q) continue;
        const double gx = (double)fx_[f] - qx, gy = (double)fy_[f] - qy;
        const double gg = gx * gx + gy * gy;
        if (gg == 0.0) continue; // co-located with q: never strictly closer
        const double dl = delta(gg) / gg;
        const double glen = std::sqrt(gg);
        for (int s = 0; s <= S; ++s) {
            ca[s] = ex_[s] * gx + ey_[s] * gy;
            xs[s] = ex_[s] * gy - ey_[s] * gx; // cross(e_s, g)
        }
        const double itol = 1e-9 * glen;
        bool relevant = dl >= 1.0;
        for (int s = 0; s < S && !relevant; ++s) {
            const bool inside = xs[s] >= -itol && -xs[s + 1] >= -itol;
            double rmin;
            if (inside) {
                rmin = 0.5 * glen;
            } else {
                const double ra = ca[s] > 0.0 ? gg / (2.0 * ca[s]) : kInf;
                const double rb = ca[s + 1] > 0.0 ? gg / (2.0 * ca[s + 1]) : kInf;
                rmin = std::min(ra, rb);
            }
            const double loose = rmin * (1.0 - dl) * (1.0 - 1e-12);
            if (loose < bound[s]) relevant = true;
        }
        if (!relevant) continue;
        out.push_back(f);
        bool changed = false;
        for (int s = 0; s < S; ++s) {
            if (!(ca[s] > 0.0 && ca[s + 1] > 0.0)) continue;
            const double rmax =
                std::max(gg / (2.0 * ca[s]), gg / (2.0 * ca[s + 1])) * (1.0 + dl) * (1.0 + 1e-12);
            std::vector<double>& h = heap[s];
            if (h.size() < k) {
                h.push_back(rmax);
                std::push_heap(h.begin(), h.end());
            } else if (rmax < h.front()) {
                std::pop_heap(h.begin(), h.end());
                h.back() = rmax;
                std::push_heap(h.begin(), h.end());
            } else {
                continue;
            }
            if (h.size() == k) {
                const double b = std::min(cap[s], h.front());
                if (b < bound[s]) {
                    bound[s] = b;
                    changed = true;
                }
            }
        }
        if (changed) {
            R = 0.0;
            for (int s = 0; s < S; ++s) R = std::max(R, bound[s]);
        }
    }
    if (st) {
        st->visited = visited;
        st->candidates = out.size();
        st->radius = R;
    }
}

unsigned filter_threads(size_t n, unsigned threads)
{
    if (threads == 0) {
        const char* env = std::getenv("RK_FILTER_THREADS");
        threads = env ? (unsigned)std::max(1, std::atoi(env)) : std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    }
    // a thread per 8 queries at least: below that the spawn costs more than it saves
    return (unsigned)std::max<size_t>(1, std::min<size_t>(threads, n / 8));
}

void filter_many(const FacilityIndex& idx, const uint32_t* q, const uint32_t* k, size_t n,
                 std::vector<std::vector<uint32_t>>& out, unsigned threads)
{
    out.resize(n);
    const unsigned t = filter_threads(n, threads);
    if (t <= 1) {
        for (size_t i = 0; i < n; ++i) idx.filter(q[i], k[i], out[i]);
        return;
    }
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t i; (i = next.fetch_add(1)) < n;) idx.filter(q[i], k[i], out[i]);
    };
    std::vector<std::thread> pool;
    for (unsigned j = 1; j < t; ++j) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

} // namespace rk

#include "rknn.h"

extern "C" rk_status rk_filter_points(const float* fx, const float* fy, size_t nf, const float* ux, const float* uy,
                                      size_t nu, uint32_t q, uint32_t k, uint32_t* cand, size_t cap, size_t* nc)
{
    if (!nc || (nf && (!fx || !fy)) || (nu && (!ux || !uy)) || (cap && !cand)) return RK_ERR_INVALID_ARG;
    if (q >= nf || k == 0) return RK_ERR_INVALID_ARG;
    double box[4] = {0.0, 0.0, -1.0, -1.0};
    if (nu) {
        box[0] = box[2] = ux[0];
        box[1] = box[3] = uy[0];
        for (size_t i = 1; i < nu; ++i) {
            box[0] = std::min(box[0], (double)ux[i]);
            box[1] = std::min(box[1], (double)uy[i]);
            box[2] = std::max(box[2], (double)ux[i]);
            box[3] = std::max(box[3], (double)uy[i]);
        }
    }
    rk::FacilityIndex idx(fx, fy, nf, box);
    std::vector<uint32_t> v;
    idx.filter(q, k, v);
    *nc = v.size();
    if (v.size() > cap) return RK_ERR_CAPACITY;
    std::copy(v.begin(), v.end(), cand);
    return RK_OK;
}
