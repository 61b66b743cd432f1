// rk_common.hpp -- layout and arithmetic shared by the host and the gfx950
// kernels of the RkNN dominance-count engine.
//
// The predicate is rknn-inf/inf.h:77-93 (a user violates bisector(q,f) when it
// lies strictly on f's side) restated as a distance comparison, so that no
// slope-intercept form is needed:
//
//   f strictly closer to u than q   <=>   m = (u-q).(f-q) - |f-q|^2/2 > 0,
//
// since |u-f|^2 - |u-q|^2 = |f-q|^2 - 2(u-q).(f-q) = -2m.  This replaces the
// rknn-rt scene (rknn-rt/rt.cpp:668-756: one mesh per bisector covering f's
// half-plane) and its any-hit counter (rknn-rt/rt.cu:95-111).
//
// FP32 fast path with a certified band.  Per query q and candidate f the
// host (or the prep kernel) forms, in binary64,
//   G = f - q,  H = |G|^2/2,  and a power of two s,
// and stores the fp32 record (gx, gy, sh) = (s*Gx, s*Gy, s*H - 1/2).  A user
// evaluates, with wx = fl32(ux - qx), wy = fl32(uy - qy),
//   t = fmaf(wx, gx, fmaf(wy, gy, -sh))   ~=  s*m + 1/2.
// Error analysis (u = 2^-24, v = 2^-53; X = s|wx*||Gx|, Y = s|wy*||Gy|,
// Z = s*H + 1/2, all over exact values):
//   |t - (s*m + 1/2)| <= 4.01u (X + Y + Z)                      (fp32 part)
// The FP64 oracle decides D64(u,f) < D64(u,q); each D64 carries a relative
// error <= 4.01v, so the oracle agrees with the sign of m whenever
//   |m| > 2.01v (D(u,f) + D(u,q)),  D(u,f) + D(u,q) <= 3 D(u,q) + 4H.
// With Wx >= |ux - qx|, Wy >= |uy - qy| over every user of the dataset,
//   E = 5u (Wx|Gx| + Wy|Gy| + H) + 4.25v (3 (Wx^2 + Wy^2) + 4H) + 2^-80
// bounds both parts, and s is the largest power of two with s*E <= 1/16.
// Then fp32 error + oracle error < 1/16 + 2.01u < 1/8 in t units, so
//   t >= 5/8  =>  the oracle says f is strictly closer      (certain)
//   t <= 3/8  =>  the oracle says f is not strictly closer  (certain)
// and only 3/8 < t < 5/8 (the band) is re-evaluated in binary64 with the
// oracle's own expression (d2_exact below).  Because t is linear in u, the
// same thresholds applied at the corner of a bounding rectangle that
// minimises (maximises) t certify every point of the rectangle.
#pragma once

#include <cstdint>
#include <cmath>

#if defined(__HIPCC__)
#define RK_HD __host__ __device__
#else
#define RK_HD
#endif

namespace rk {

constexpr int kTile = 64;            // users per leaf tile: one wavefront
constexpr int kFanout = 64;          // children per tree node: one wavefront
constexpr int kMaxLevels = 8;        // 64^8 tiles: far beyond any HBM
constexpr float kCertainHi = 0.625f; // t >= 5/8: certainly closer
constexpr float kCertainLo = 0.375f; // t <= 3/8: certainly not closer

struct alignas(16) CandRec {
    float gx, gy, sh, pad;
};

// Per-query bound on |u - q| per axis over every user (and so over every
// bounding-rectangle corner), from the users' bounding box.  The 2^-20
// slack covers the binary64 subtraction.
RK_HD inline void axis_bounds(double minx, double miny, double maxx, double maxy, float qx, float qy,
                              double* wx, double* wy)
{
    double ax = std::fabs(minx - (double)qx), bx = std::fabs(maxx - (double)qx);
    double ay = std::fabs(miny - (double)qy), by = std::fabs(maxy - (double)qy);
    *wx = (ax > bx ? ax : bx) * (1.0 + 0x1p-20);
    *wy = (ay > by ? ay : by) * (1.0 + 0x1p-20);
}

// The fp32 record of candidate f for query q (see the header comment).
RK_HD inline CandRec make_record(float qx, float qy, float fx, float fy, double wx, double wy)
{
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const double gx = (double)fx - (double)qx;
    const double gy = (double)fy - (double)qy;
    const double gx2 = gx * gx;
    const double gy2 = gy * gy;
    const double h = 0.5 * (gx2 + gy2);
    const double fx32 = 5.0 * 0x1p-24 * (wx * std::fabs(gx) + wy * std::fabs(gy) + h);
    const double fx64 = 4.25 * 0x1p-53 * (3.0 * (wx * wx + wy * wy) + 4.0 * h);
    const double e = (fx32 + fx64 + 0x1p-80) * 1.0625; // slack for E's own rounding
    int ex = 0;
    (void)std::frexp(e, &ex);                 // e = m 2^ex, m in [0.5, 1)
    const double s = std::ldexp(1.0, -(ex + 4)); // s*16e = m < 1
    CandRec r;
    r.gx = (float)(s * gx);
    r.gy = (float)(s * gy);
    r.sh = (float)(s * h - 0.5);
    r.pad = 0.0f;
    return r;
}

// Squared distance exactly as the oracle evaluates it: binary64 differences
// of the fp32 coordinates, two products, one sum, no contraction.
RK_HD inline double d2_exact(float ax, float ay, float bx, float by)
{
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const double dx = (double)ax - (double)bx;
    const double dy = (double)ay - (double)by;
    const double x2 = dx * dx;
    const double y2 = dy * dy;
    return x2 + y2;
}

// Device view of a resident dataset (users Hilbert-sorted, padded to whole
// tiles by repeating the last user, which leaves every MBR unchanged).
struct DsDev {
    const float* ux;
    const float* uy;
    const int32_t* uid;
    const float4* mbr[kMaxLevels]; // mbr[0]: tiles; mbr[l]: groups of 64 level-(l-1) nodes
    const float4* gmbr4;           // groups of 4 consecutive tiles (count-all)
    const float4* gmbr6;           // groups of 6 consecutive tiles (count-all)
    const float4* gmbr8;           // groups of 8 consecutive tiles (count-all)
    uint32_t nlev[kMaxLevels];     // node count per level
    uint32_t n_levels;             // levels present; nlev[n_levels-1] <= 64
    uint32_t nu;                   // real users (lanes >= nu are padding)
    uint32_t n_tiles;
};

// One query as the kernels see it.
struct QDev {
    float qx, qy;
    uint32_t k;
    uint32_t nc;      // candidate records
    uint64_t off;     // into the batch's rec[] / orig[] arrays
    uint32_t trivial; // k > nc: every user qualifies, no traversal
    uint32_t pad;
};

} // namespace rk
