// rk_dfilter.hip -- the influence-zone filter on the device.
//
// The same sector-bound filter as the host's FacilityIndex::filter
// (rk_filter.cpp; it replaces the bisector generation of
// rknn-inf/inf.cpp:201-585), run by one wave per query over a facility grid
// kept in HBM, so a batch of (q, k) requests goes from query to result ids
// without leaving the device.
//
//  * Facilities are bucketed in a uniform grid (CSR: cell start offsets,
//    coordinates and original indices sorted by cell).  The wave visits the
//    cells in rings of growing Chebyshev radius around q's cell; a ring r >= 2
//    is at least (r - 1) cell sides from q, so the walk stops once half that
//    distance (shrunk by the oracle's rounding allowance) reaches the largest
//    sector bound, exactly the host stream's stopping rule.
//  * A ring's facilities are tested 64 at a time (lane = facility) against
//    the sector bounds of the moment: a facility is relevant when its
//    half-plane, loosened by the oracle's error, reaches inside some sector's
//    bound.  The test walks the sectors outward from the one holding g = f - q,
//    where the half-plane's entry distance is smallest and grows on both
//    sides, so it stops early.
//  * The batch's relevant facilities join the candidate list (compacted by a
//    ballot into LDS) and then update the sectors (lane = sector, each lane
//    walking the batch's relevant facilities on its own): every sector keeps
//    the k smallest exit distances (tightened by the oracle's error, stored
//    as fp32 rounded up) in a sorted register list, and once it holds k its
//    bound is the k-th.  The largest bound over the sectors is refreshed once
//    per batch (a stale, larger value only stops the walk later).
//
// The visiting order differs from the host's nearest-first stream, and the
// relevance test of a batch uses the bounds from before the batch; neither
// changes the sufficiency argument (a pruned facility is pruned against
// bounds backed by k candidates already in the list, and bounds only
// shrink), so the candidate list may differ from the host's but every
// count over it gives the oracle's result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "rk_dfilter.hpp"
#include "rk_internal.hpp"
#include "rknn.h"

namespace rk {

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

__device__ __forceinline__ double wave_max_d(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// exclusive prefix sum over the wave; *total = the sum over all lanes
__device__ __forceinline__ uint32_t wave_scan_excl(uint32_t v, uint32_t* total)
{
    const uint32_t lane = __lane_id();
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    *total = (uint32_t)__shfl((int)x, 63);
    return x - v;
}

// Cell i of the ring of Chebyshev radius r around (cx, cy): 8r cells, four
// sides of 2r cells each (r = 0: the centre).
__device__ __forceinline__ void ring_cell(int cx, int cy, int r, uint32_t i, int& x, int& y)
{
    if (r == 0) {
        x = cx;
        y = cy;
        return;
    }
    const int side = (int)i / (2 * r), pos = (int)i % (2 * r);
    switch (side) {
    case 0: x = cx - r + pos; y = cy - r; break;
    case 1: x = cx + r; y = cy - r + pos; break;
    case 2: x = cx + r - pos; y = cy + r; break;
    default: x = cx - r; y = cy + r - pos; break;
    }
}

// The binary64 oracle's error on D(u,f) - D(u,q) over the users' box,
// doubled (as in the host filter).
__device__ __forceinline__ double df_delta(double rbox2, double gg)
{
    return 8.02 * 0x1p-53 * (3.0 * rbox2 + 2.0 * gg) + 0x1p-100;
}

// The loosened smallest entry distance l of half-plane g into sector s, as
// a pair (lhs, unit) with  l < B  <=>  lhs < unit * B  (no division):
//  * g's direction inside the sector: l = |g|/2 * shrink, so (hl, 1);
//  * else l = |g|^2 / (2 c) * shrink over the sector's ray with the larger
//    c = e.g > 0, so (|g|^2 * shrink, 2 c); no ray with c > 0: never.
// shrink = (1 - dl)(1 - 1e-12) loosens by the oracle's error; the few ulps
// of the multiplied-out form lie far inside its 1e-12.
struct Entry {
    double lhs, unit;
};

__device__ __forceinline__ Entry df_entry(const double* ex, const double* ey, int s, double gx, double gy, double hl,
                                          double ggs, double itol)
{
    const double xa = ex[s] * gy - ey[s] * gx, xb = ex[s + 1] * gy - ey[s + 1] * gx;
    if (xa >= -itol && -xb >= -itol) return {hl, 1.0};
    const double ca = ex[s] * gx + ey[s] * gy, cb = ex[s + 1] * gx + ey[s + 1] * gy;
    const double m = fmax(ca, cb);
    return m > 0.0 ? Entry{ggs, 2.0 * m} : Entry{1.0, 0.0};
}

__device__ __forceinline__ bool df_inside(const double* ex, const double* ey, int s, double gx, double gy, double itol)
{
    const double xa = ex[s] * gy - ey[s] * gx, xb = ex[s + 1] * gy - ey[s + 1] * gx;
    return xa >= -itol && -xb >= -itol;
}

// Is facility g relevant against the bounds?  R is the largest bound.
__device__ bool df_relevant(const double* ex, const double* ey, const double* bound, double gx, double gy, double gg,
                            double dl, double R)
{
    if (dl >= 1.0) return true;
    // every sector's entry is at least |g|/2 (squared test first: no root)
    const double shrink = (1.0 - dl) * (1.0 - 1e-12);
    if (0.25 * gg * shrink * shrink >= R * R * (1.0 + 1e-9)) return false;
    const double glen = sqrt(gg);
    const double hl = 0.5 * glen * shrink, ggs = gg * shrink;
    if (hl >= R) return false;
    const double itol = 1e-9 * glen;
    float ang = atan2f((float)gy, (float)gx);
    if (ang < 0.f) ang += (float)kTwoPi;
    int s0 = min(max((int)(ang * (float)(kDfSectors / kTwoPi)), 0), kDfSectors - 1);
    if (!df_inside(ex, ey, s0, gx, gy, itol)) {
        const int sl = (s0 + kDfSectors - 1) % kDfSectors, sr = (s0 + 1) % kDfSectors;
        if (df_inside(ex, ey, sl, gx, gy, itol)) {
            s0 = sl;
        } else if (df_inside(ex, ey, sr, gx, gy, itol)) {
            s0 = sr;
        } else { // no sector found near the estimate: test them all
            for (int s = 0; s < kDfSectors; ++s) {
                const Entry e = df_entry(ex, ey, s, gx, gy, hl, ggs, itol);
                if (e.lhs < e.unit * bound[s]) return true;
            }
            return false;
        }
    }
    // from the sector holding g outwards the entry distance only grows
    const double stop = R * (1.0 + 1e-9);
    bool right = true, left = true;
    for (int d = 0; d < kDfSectors / 2 && (right || left); ++d) {
        if (right) {
            const int s = (s0 + d) % kDfSectors;
            const Entry e = df_entry(ex, ey, s, gx, gy, hl, ggs, itol);
            if (e.lhs < e.unit * bound[s]) return true;
            right = e.lhs < e.unit * stop;
        }
        if (left) {
            const int s = (s0 + kDfSectors - 1 - d) % kDfSectors;
            const Entry e = df_entry(ex, ey, s, gx, gy, hl, ggs, itol);
            if (e.lhs < e.unit * bound[s]) return true;
            left = e.lhs < e.unit * stop;
        }
    }
    return false;
}

// Each sector lane keeps the K smallest exit distances it has seen, sorted,
// in registers: K = 16, 32 or 64, the smallest that holds k.  One wave runs
// one query, a wave to a SIMD for a batch of about a thousand queries, so
// registers are not what limits the filter: the dependent latency of its
// instructions is.  An insertion is K independent compare/select pairs.
//
// Inserts v into h[0 .. K), sorted ascending (h[i] becomes the larger of
// h[i-1] and min(h[i], v): every new value depends only on old ones).
template <int K>
__device__ __forceinline__ void reg_insert(float* h, float v)
{
#pragma unroll
    for (int i = K - 1; i > 0; --i) h[i] = v < h[i - 1] ? h[i - 1] : fminf(h[i], v);
    h[0] = fminf(h[0], v);
}

// h[k - 1] of the sorted (non-negative) list, as the largest of its first k:
// a reduction, so the array is never indexed at run time (which would move
// it to scratch memory)
template <int K>
__device__ __forceinline__ float reg_kth(const float* h, uint32_t k)
{
    float t = 0.f;
#pragma unroll
    for (int i = 0; i < K; ++i) t = fmaxf(t, (uint32_t)i < k ? h[i] : 0.f);
    return t;
}

// fp32 value >= v
__device__ __forceinline__ float round_up_f(double v)
{
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    return f;
}

// Exit distance of relevant facility e from the sector (ax, ay)..(bx, by),
//   rmax = |g|^2 (1 + dl)(1 + 1e-12) / (2 min(ca, cb)),
// tightened by the oracle's error, as an fp32 value >= rmax; infinite when
// the half-plane does not cross both of the sector's rays.  e.z holds the
// numerator, already raised to an fp32 value >= it (once per facility).
// The quotient is taken in fp32 with directed margins: the denominator is
// lowered by 2^-22 before its conversion (which rounds by <= 2^-24), the
// reciprocal (<= 1 ulp) and the product (<= 1/2 ulp) are covered by the
// final factor 1 + 2^-21 (its own rounding <= 1/2 ulp).  A denominator
// below fp32's range gives an infinite value, and one of 2^100 or more (its
// reciprocal near the flushed range) is not used: an infinite exit distance
// only leaves the sector's bound looser, which is always safe.
__device__ __forceinline__ float exit_dist(const double4& e, double ax, double ay, double bx, double by)
{
    const double ca = ax * e.x + ay * e.y, cb = bx * e.x + by * e.y;
    const float den = (float)(2.0 * fmin(ca, cb) * (1.0 - 0x1p-22));
    const bool use = ca > 0.0 && cb > 0.0 && den < 0x1p100f;
    const float v = (float)e.z * __builtin_amdgcn_rcpf(den) * (1.0f + 0x1p-21f);
    return use ? v : INFINITY;
}

struct DfShared {
    double bound[kDfSectors];
    double ex[kDfSectors + 1], ey[kDfSectors + 1];
    uint32_t st[64], pre[64];
    double4 rel[64]; // the batch's relevant facilities (exit_dist's operands)
};

// The candidate list of one query (q, k <= K) into out[0 .. cap), its length
// returned (it may exceed cap: overflow).  One wave; lane = sector.
template <int K>
__device__ uint32_t filter_query(const FacGridDev& g, DfShared& sh, uint32_t q, uint32_t k, uint32_t* out,
                                 uint32_t cap)
{
    const uint32_t lane = threadIdx.x;
    const double ax = sh.ex[lane], ay = sh.ey[lane], bxs = sh.ex[lane + 1], bys = sh.ey[lane + 1];
    const double bx0 = g.box[0], by0 = g.box[1], bx1 = g.box[2], by1 = g.box[3];
    const int G = (int)g.G;
    const double qx = g.ofx[q], qy = g.ofy[q];
    uint32_t n = 0;
    // farthest squared distance from q to the users' box
    double rbox2 = 0.0;
    const double cx4[4] = {bx0, bx1, bx0, bx1}, cy4[4] = {by0, by0, by1, by1};
    for (int i = 0; i < 4; ++i) {
        const double dx = cx4[i] - qx, dy = cy4[i] - qy;
        rbox2 = fmax(rbox2, dx * dx + dy * dy);
    }
    // this lane's sector: the farthest point of the users' box inside it
    double best = -1.0;
    {
        const double tol = 1e-12;
        if (qx >= bx0 && qx <= bx1 && qy >= by0 && qy <= by1) best = 0.0;
        for (int i = 0; i < 4; ++i) {
            const double dx = cx4[i] - qx, dy = cy4[i] - qy;
            const double sc = tol * (fabs(dx) + fabs(dy));
            if (ax * dy - ay * dx >= -sc && dx * bys - dy * bxs >= -sc) best = fmax(best, dx * dx + dy * dy);
        }
        const double span = (bx1 - bx0) + (by1 - by0) + 1.0;
        for (int e = 0; e < 2; ++e) {
            const double ux = e ? bxs : ax, uy = e ? bys : ay;
            if (ux != 0.0) {
                for (int w = 0; w < 2; ++w) {
                    const double X = w ? bx1 : bx0;
                    const double t = (X - qx) / ux;
                    const double y = qy + t * uy;
                    if (t >= 0.0 && y >= by0 - tol * span && y <= by1 + tol * span) best = fmax(best, t * t);
                }
            }
            if (uy != 0.0) {
                for (int w = 0; w < 2; ++w) {
                    const double Y = w ? by1 : by0;
                    const double t = (Y - qy) / uy;
                    const double x = qx + t * ux;
                    if (t >= 0.0 && x >= bx0 - tol * span && x <= bx1 + tol * span) best = fmax(best, t * t);
                }
            }
        }
    }
    const double capv = best < 0.0 ? 0.0 : sqrt(best) * (1.0 + 1e-9) + 1e-9;
    double bound = capv;
    sh.bound[lane] = bound;
    double R = wave_max_d(bound);
    float h[K];
#pragma unroll
    for (int i = 0; i < K; ++i) h[i] = INFINITY;
    __syncthreads();
    const bool have_users = bx0 <= bx1 && by0 <= by1;
    const int cx = min(max((int)floor((qx - g.x0) / g.cs), 0), G - 1);
    const int cy = min(max((int)floor((qy - g.y0) / g.cs), 0), G - 1);
    for (int r = 0; have_users && k > 0; ++r) {
        if (r >= 2) {
            const double dmin = (double)(r - 1) * g.cs * (1.0 - 1e-9);
            const double key = dmin * dmin;
            const double dl = df_delta(rbox2, key) / key;
            const double lower = dl >= 1.0 ? 0.0 : 0.5 * dmin * (1.0 - dl) * (1.0 - 1e-12);
            if (lower >= R) break;
        }
        if (r > 0 && cx - r < 0 && cx + r >= G && cy - r < 0 && cy + r >= G) break; // past every cell
        const uint32_t ncells = r == 0 ? 1u : 8u * (uint32_t)r;
        for (uint32_t c0 = 0; c0 < ncells; c0 += 64) {
            uint32_t st = 0, cnt = 0;
            if (c0 + lane < ncells) {
                int x, y;
                ring_cell(cx, cy, r, c0 + lane, x, y);
                if (x >= 0 && x < G && y >= 0 && y < G) {
                    const uint32_t c = (uint32_t)y * (uint32_t)G + (uint32_t)x;
                    st = g.start[c];
                    cnt = g.start[c + 1] - st;
                }
            }
            uint32_t total = 0;
            const uint32_t pre = wave_scan_excl(cnt, &total);
            sh.st[lane] = st;
            sh.pre[lane] = pre;
            __syncthreads();
            for (uint32_t j0 = 0; j0 < total; j0 += 64) {
                const uint32_t j = j0 + lane;
                bool rel = false;
                double gx = 0.0, gy = 0.0, gg = 0.0, dl = 0.0;
                uint32_t f = 0;
                if (j < total) {
                    int lo = 0, hi = 63; // the last cell whose prefix is <= j (it holds slot j)
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (sh.pre[mid] <= j) lo = mid;
                        else hi = mid - 1;
                    }
                    const uint32_t slot = sh.st[lo] + (j - sh.pre[lo]);
                    f = g.fid[slot];
                    gx = (double)g.fx[slot] - qx;
                    gy = (double)g.fy[slot] - qy;
                    gg = gx * gx + gy * gy;
                    if (f != q && gg != 0.0) { // co-located with q: never strictly closer
                        dl = df_delta(rbox2, gg) / gg;
                        rel = df_relevant(sh.ex, sh.ey, sh.bound, gx, gy, gg, dl, R);
                    }
                }
                // candidates, in lane order, and their parameters for the sector lanes
                const uint64_t rm = __builtin_amdgcn_ballot_w64(rel);
                const uint32_t nrel = (uint32_t)__popcll(rm);
                if (rel) {
                    const uint32_t rk = __builtin_amdgcn_mbcnt_hi((uint32_t)(rm >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)rm, 0u));
                    // (gx, gy, fp32 value >= |g|^2 (1 + dl)(1 + 1e-12), unused)
                    sh.rel[rk] = make_double4(gx, gy, (double)round_up_f(gg * (1.0 + dl) * (1.0 + 1e-12)), 0.0);
                    if (n + rk < cap) out[n + rk] = f;
                }
                n += nrel;
                __syncthreads();
                // lane = sector: the relevant facilities' exit distances from
                // this sector, four at a time (independent chains), then
                // inserted; an insertion no lane needs is skipped
                for (uint32_t i0 = 0; i0 < nrel; i0 += 4) {
                    float v[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        v[u] = i0 + u < nrel ? exit_dist(sh.rel[i0 + u], ax, ay, bxs, bys) : INFINITY;
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (__builtin_amdgcn_ballot_w64(v[u] < h[K - 1]) != 0) reg_insert<K>(h, v[u]);
                }
                if (nrel) {
                    const float kth = reg_kth<K>(h, k);
                    if (kth < INFINITY) { // the sector has k exits
                        const double bnd = fmin(capv, (double)kth);
                        if (bnd < bound) {
                            bound = bnd;
                            sh.bound[lane] = bnd;
                        }
                    }
                    R = wave_max_d(bound);
                }
                __syncthreads(); // sh.bound is read by the next batch's relevance tests
            }
            __syncthreads(); // sh.st / sh.pre are rewritten for the next chunk
        }
    }
    return n;
}

// One wave per query.  K = 16, 32 or 64: the kernel runs the batch's
// queries whose k takes lists of K entries and skips the others, with the
// registers of its own list size (118, 134 and 166 VGPRs), so a batch of
// small k leaves room for the traversal kernels of the steps beside it.
// K = 0: every query, with the list size its k takes (the registers of the
// largest), for batches that mix list sizes -- one launch per size would
// run the sizes' slowest queries one after another.
template <int K>
__global__ __launch_bounds__(64) void k_dfilter(FacGridDev g, const uint32_t* __restrict__ qs,
                                                const uint32_t* __restrict__ ks, uint32_t nq,
                                                uint32_t* __restrict__ cand, uint32_t cap,
                                                uint32_t* __restrict__ ncand)
{
    __shared__ DfShared sh;
    const uint32_t lane = threadIdx.x;
    sh.ex[lane] = g.ex[lane];
    sh.ey[lane] = g.ey[lane];
    if (lane == 0) {
        sh.ex[kDfSectors] = g.ex[kDfSectors];
        sh.ey[kDfSectors] = g.ey[kDfSectors];
    }
    __syncthreads();
    for (uint32_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
        const uint32_t k = min(ks[qi], kDfMaxK);
        if (K != 0 && dfilter_list(k) != (uint32_t)K) continue; // another kernel's query (uniform)
        const uint32_t q = qs[qi];
        uint32_t* out = cand + (uint64_t)qi * cap;
        uint32_t n;
        if (K == 16 || (K == 0 && k <= 16))
            n = filter_query<16>(g, sh, q, k, out, cap);
        else if (K == 32 || (K == 0 && k <= 32))
            n = filter_query<32>(g, sh, q, k, out, cap);
        else
            n = filter_query<64>(g, sh, q, k, out, cap);
        if (lane == 0) ncand[qi] = n;
        __syncthreads();
    }
}

// Per-query device view of a device-filtered batch: QDev from the filter's
// counts (a count above the segment capacity marks the query overflowed).
__global__ void k_dfilter_setup(FacGridDev g, const uint32_t* __restrict__ qs, const uint32_t* __restrict__ ks,
                                uint32_t nq, const uint32_t* __restrict__ ncand, uint32_t cap, uint32_t nu,
                                QDev* __restrict__ qd, double2* __restrict__ wb)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
        const uint32_t q = qs[i];
        QDev d;
        d.qx = g.ofx[q];
        d.qy = g.ofy[q];
        d.k = ks[i];
        d.nc = min(ncand[i], cap);
        d.off = (uint64_t)i * cap;
        d.trivial = (d.k > d.nc || nu == 0) ? 1u : 0u;
        d.pad = 0;
        qd[i] = d;
        double wx = 0.0, wy = 0.0;
        if (nu) axis_bounds(g.box[0], g.box[1], g.box[2], g.box[3], d.qx, d.qy, &wx, &wy);
        wb[i] = make_double2(wx, wy);
    }
}

} // namespace

rk_status build_facility_grid(rk_dataset* ds)
{
    rk_ctx* c = ds->ctx;
    const size_t nf = ds->nf;
    const float* fx = ds->hfx.data();
    const float* fy = ds->hfy.data();
    FacGridDev& v = ds->fgrid;
    v = FacGridDev{};
    double x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    if (nf) {
        x0 = x1 = fx[0];
        y0 = y1 = fy[0];
        for (size_t i = 1; i < nf; ++i) {
            x0 = std::min(x0, (double)fx[i]);
            y0 = std::min(y0, (double)fy[i]);
            x1 = std::max(x1, (double)fx[i]);
            y1 = std::max(y1, (double)fy[i]);
        }
    }
    // about two facilities per cell (4, 8 and 16 measured slower: fewer
    // rings, but more facilities tested against looser bounds)
    const uint32_t G = (uint32_t)std::max<double>(1.0, std::ceil(std::sqrt((double)nf / 2.0)));
    const double ext = std::max(x1 - x0, y1 - y0);
    const double cs = ext > 0 ? ext / G : 1.0;
    std::vector<uint32_t> cell(nf), start((size_t)G * G + 1, 0);
    for (size_t i = 0; i < nf; ++i) {
        const int ix = std::min(std::max((int)std::floor(((double)fx[i] - x0) / cs), 0), (int)G - 1);
        const int iy = std::min(std::max((int)std::floor(((double)fy[i] - y0) / cs), 0), (int)G - 1);
        cell[i] = (uint32_t)iy * G + (uint32_t)ix;
        ++start[cell[i] + 1];
    }
    for (size_t i = 1; i < start.size(); ++i) start[i] += start[i - 1];
    std::vector<uint32_t> pos(start.begin(), start.end() - 1), gid(nf);
    std::vector<float> gx(nf), gy(nf);
    for (size_t i = 0; i < nf; ++i) {
        const uint32_t p = pos[cell[i]]++;
        gid[p] = (uint32_t)i;
        gx[p] = fx[i];
        gy[p] = fy[i];
    }
    double ex[kDfSectors + 1], ey[kDfSectors + 1];
    for (int s = 0; s < kDfSectors; ++s) {
        ex[s] = std::cos(kTwoPi * s / kDfSectors);
        ey[s] = std::sin(kTwoPi * s / kDfSectors);
    }
    ex[kDfSectors] = ex[0];
    ey[kDfSectors] = ey[0];
    hipStream_t st = c->stream;
    RK_HIP(c, ds->g_start.reserve(start.size()));
    RK_HIP(c, ds->g_fid.reserve(std::max<size_t>(1, nf)));
    RK_HIP(c, ds->g_fx.reserve(std::max<size_t>(1, nf)));
    RK_HIP(c, ds->g_fy.reserve(std::max<size_t>(1, nf)));
    RK_HIP(c, ds->g_dir.reserve(2 * (kDfSectors + 1)));
    RK_HIP(c, hipMemcpyAsync(ds->g_start.p, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    if (nf) {
        RK_HIP(c, hipMemcpyAsync(ds->g_fid.p, gid.data(), nf * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        RK_HIP(c, hipMemcpyAsync(ds->g_fx.p, gx.data(), nf * sizeof(float), hipMemcpyHostToDevice, st));
        RK_HIP(c, hipMemcpyAsync(ds->g_fy.p, gy.data(), nf * sizeof(float), hipMemcpyHostToDevice, st));
    }
    RK_HIP(c, hipMemcpyAsync(ds->g_dir.p, ex, sizeof ex, hipMemcpyHostToDevice, st));
    RK_HIP(c, hipMemcpyAsync(ds->g_dir.p + kDfSectors + 1, ey, sizeof ey, hipMemcpyHostToDevice, st));
    RK_HIP(c, hipStreamSynchronize(st)); // the host vectors die here
    v.fx = ds->g_fx.p;
    v.fy = ds->g_fy.p;
    v.fid = ds->g_fid.p;
    v.start = ds->g_start.p;
    v.ofx = ds->fx.p;
    v.ofy = ds->fy.p;
    v.ex = ds->g_dir.p;
    v.ey = ds->g_dir.p + kDfSectors + 1;
    v.x0 = x0;
    v.y0 = y0;
    v.cs = cs;
    v.G = G;
    v.nf = (uint32_t)nf;
    for (int i = 0; i < 4; ++i) v.box[i] = ds->ubox[i];
    return RK_OK;
}

uint32_t dfilter_lists(const uint32_t* k, size_t n)
{
    uint32_t m = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t l = dfilter_list(std::min(k[i], kDfMaxK));
        m |= l == 16 ? 1u : l == 32 ? 2u : 4u;
    }
    return m;
}

void launch_dfilter(const FacGridDev& g, const uint32_t* q, const uint32_t* k, uint32_t nq, uint32_t lists,
                    uint32_t* cand, uint32_t cap, uint32_t* ncand, hipStream_t st)
{
    if (!nq) return;
    const uint32_t grid = nq < 65535 ? nq : 65535;
    if (lists == 1u)
        k_dfilter<16><<<grid, 64, 0, st>>>(g, q, k, nq, cand, cap, ncand);
    else if (lists == 2u)
        k_dfilter<32><<<grid, 64, 0, st>>>(g, q, k, nq, cand, cap, ncand);
    else if (lists == 4u)
        k_dfilter<64><<<grid, 64, 0, st>>>(g, q, k, nq, cand, cap, ncand);
    else // mixed list sizes (or unknown): one launch
        k_dfilter<0><<<grid, 64, 0, st>>>(g, q, k, nq, cand, cap, ncand);
}

void launch_dfilter_setup(const FacGridDev& g, const uint32_t* q, const uint32_t* k, uint32_t nq,
                          const uint32_t* ncand, uint32_t cap, uint32_t nu, QDev* qd, double2* wb, hipStream_t st)
{
    if (!nq) return;
    const uint32_t blocks = std::min<uint32_t>((nq + 255) / 256, 1024);
    k_dfilter_setup<<<blocks, 256, 0, st>>>(g, q, k, nq, ncand, cap, nu, qd, wb);
}

} // namespace rk

using namespace rk;

extern "C" rk_status rk_filter_device(rk_dataset* ds, const uint32_t* q, const uint32_t* k, size_t nq, uint32_t* cand,
                                      size_t cap_per_query, uint64_t* nc)
{
    if (!ds || !nc || (nq && (!q || !k)) || (nq && cap_per_query && !cand)) return RK_ERR_INVALID_ARG;
    rk_ctx* c = ds->ctx;
    if (nq > (1u << 24)) return fail(c, RK_ERR_INVALID_ARG, "too many queries in one call");
    if (cap_per_query > 0xffffffffull) return fail(c, RK_ERR_INVALID_ARG, "capacity per query too large");
    for (size_t i = 0; i < nq; ++i) {
        if (q[i] >= ds->nf) return fail(c, RK_ERR_INVALID_ARG, "query index q must be < number of facilities");
        if (k[i] == 0) return fail(c, RK_ERR_INVALID_ARG, "k must be a positive integer");
        if (k[i] > kDfMaxK) return fail(c, RK_ERR_INVALID_ARG, "the device filter takes k <= 64");
    }
    if (nq == 0) return RK_OK;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const uint32_t capq = (uint32_t)std::max<size_t>(1, cap_per_query);
    DevBuf<uint32_t> dq, dk, dc, dn;
    RK_HIP(c, dq.reserve(nq));
    RK_HIP(c, dk.reserve(nq));
    RK_HIP(c, dn.reserve(nq));
    RK_HIP(c, dc.reserve((uint64_t)nq * capq));
    RK_HIP(c, hipMemcpyAsync(dq.p, q, nq * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    RK_HIP(c, hipMemcpyAsync(dk.p, k, nq * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    launch_dfilter(ds->fgrid, dq.p, dk.p, (uint32_t)nq, dfilter_lists(k, nq), dc.p, capq, dn.p, st);
    RK_HIP(c, hipGetLastError());
    std::vector<uint32_t> cnt(nq);
    RK_HIP(c, hipMemcpyAsync(cnt.data(), dn.p, nq * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    if (cap_per_query)
        RK_HIP(c, hipMemcpyAsync(cand, dc.p, (uint64_t)nq * capq * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    RK_HIP(c, hipStreamSynchronize(st));
    bool over = false;
    for (size_t i = 0; i < nq; ++i) {
        nc[i] = cnt[i];
        over |= cnt[i] > cap_per_query;
    }
    return over ? fail(c, RK_ERR_CAPACITY, "a query's candidate list exceeds cap_per_query") : RK_OK;
}
