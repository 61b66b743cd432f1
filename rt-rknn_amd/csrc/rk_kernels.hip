// rk_kernels.hip -- gfx950 kernels of the RkNN dominance-count engine.
//
// Wave64 throughout: a leaf tile is 64 Hilbert-adjacent users (one per lane),
// a tree node has 64 children (one per lane), and candidate records are
// staged 64 at a time in a wave-private LDS slot and read back as
// broadcasts.  Blocks are 256 threads = 4 independent waves; no kernel uses
// a block barrier inside a data-dependent loop.
//
// Reference counterparts: check_users_kernel (rknn-inf-gpu/inf_kernel.cu:41-66,
// one thread per user over all bisectors, early exit at k),
// findKNearestSimpleKernel (rknn-naive-gpu/rknn_kernel.cu:119-169), and the
// rknn-rt raygen/any-hit pair (rknn-rt/rt.cu:46-81,95-111).
#include <hip/hip_runtime.h>
#include <algorithm>

#include "rk_common.hpp"
#include "rk_kernels.hpp"

namespace rk {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Lane mask of a predicate (the wave64 ballot straight from the compare,
// without materialising 0/1 in a vector register first).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// 3/8 < t < 5/8 exactly: t - 1/2 is exact for t in [1/4, 1] (Sterbenz), and
// outside that interval |fl(t - 1/2)| >= 1/4 by monotone rounding.
__device__ __forceinline__ bool in_band(float t) { return fabsf(t - 0.5f) < 0.125f; }

__device__ __forceinline__ uint32_t rank_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Values every lane holds identically, moved to scalar registers so that the
// loops and loads depending on them compile to scalar branches and loads.
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform64(uint64_t v)
{
    return ((uint64_t)uniform((uint32_t)(v >> 32)) << 32) | uniform((uint32_t)v);
}
__device__ __forceinline__ float uniformf(float v) { return __int_as_float((int)uniform((uint32_t)__float_as_int(v))); }
__device__ __forceinline__ QDev uniform_q(const QDev& q)
{
    QDev r;
    r.qx = uniformf(q.qx);
    r.qy = uniformf(q.qy);
    r.k = uniform(q.k);
    r.nc = uniform(q.nc);
    r.off = uniform64(q.off);
    r.trivial = uniform(q.trivial);
    r.pad = 0;
    return r;
}

struct LeafOut {
    uint64_t keep;     // lanes (users) of the tile in R(q,k)
    uint32_t fp64;     // (user, candidate) tests resolved in binary64
};

// Candidate records of one query live in a block-shared LDS array when there
// are at most N of them (built or loaded once per block, then read as
// broadcasts by every wave).  Queries with more candidates stream 64-record
// chunks from global memory through a wave-private LDS slot (WaveLds); the
// two layouts share the same LDS bytes.
struct WaveLds {
    CandRec rec[64];
    float2 orig[64];
};

template <uint32_t N>
struct SharedRecs {
    CandRec rec[N];
    float2 orig[N];
};

// The batched form's shared records: the facility index in each record's
// spare word instead of an orig[] array (8 bytes of LDS less per record).
template <uint32_t N>
struct SharedRecsIdx {
    CandRec rec[N];
};

// One query against one tile of 64 users.  Phase A puts candidates on lanes
// and classifies each against the tile's bounding rectangle m: certainly
// covering the whole tile (counted once for every user), certainly missing
// it, or crossing it.  Phase B puts users on lanes and evaluates only the
// crossing candidates, four per pass so that their LDS reads overlap; a
// (user, candidate) test whose fp32 value lands in the band is settled on
// the spot by the oracle's binary64 expression, so the count is exact when
// the loop ends.
//   kShared: rec/orig are the block-shared LDS arrays (else global, staged
//            64 at a time in L)
//   kPre:    the users' coordinates are given (pux, puy); otherwise they are
//            loaded on the first crossing candidate, so culled tiles cost no
//            user traffic
//   kIdx:    (with kShared) the records carry their facility index in the
//            spare word, and a band test reads the facility's coordinates
//            from fxg/fyg instead of orig[]
template <bool kShared, bool kPre, bool kIdx = false>
__device__ LeafOut leaf_tile(const DsDev& ds, float qx, float qy, uint32_t k, uint32_t nc,
                             const CandRec* __restrict__ rec, const float2* __restrict__ orig, uint32_t tile,
                             float4 m, float pux, float puy, WaveLds* L, const float* __restrict__ fxg = nullptr,
                             const float* __restrict__ fyg = nullptr)
{
    static_assert(!kIdx || kShared, "indexed records are the shared-LDS layout");
    const uint32_t lane = lane_id();
    k = uniform(k);
    nc = uniform(nc);
    tile = uniform(tile);
    const uint32_t u = tile * kTile + lane;
    const bool valid = u < ds.nu;
    float ux = pux, uy = puy, wx = 0.f, wy = 0.f;
    double dq = 0.0;
    bool loaded = false; // wave-uniform
    const float lx = m.x - qx, ly = m.y - qy, hx = m.z - qx, hy = m.w - qy;

    uint32_t base = 0; // wave-uniform: candidates covering the whole tile
    uint32_t cnt = 0;  // per user: crossing candidates closer (exact)
    uint32_t fp64 = 0; // wave-uniform
    bool dead = false; // every user already has >= k
    for (uint32_t c0 = 0; c0 < nc; c0 += 64) {
        const uint32_t j = c0 + lane;
        const bool have = j < nc;
        CandRec r = {0.f, 0.f, 0.f, 0.f};
        float2 o = make_float2(0.f, 0.f);
        if (have) {
            r = rec[j];
            if (!kShared) o = orig[j];
        }
        const float tlo = fmaf(r.gx >= 0.f ? lx : hx, r.gx, fmaf(r.gy >= 0.f ? ly : hy, r.gy, -r.sh));
        const float thi = fmaf(r.gx >= 0.f ? hx : lx, r.gx, fmaf(r.gy >= 0.f ? hy : ly, r.gy, -r.sh));
        const bool full = have && tlo >= kCertainHi;
        const bool cross = have && !full && thi > kCertainLo;
        base += (uint32_t)__popcll(ballot(full));
        if (base >= k) {
            dead = true;
            break;
        }
        uint64_t cm = ballot(cross);
        if (cm) {
            if (!loaded) { // (arrays are padded to whole tiles)
                if (!kPre) {
                    ux = ds.ux[u];
                    uy = ds.uy[u];
                }
                wx = ux - qx;
                wy = uy - qy;
                dq = d2_exact(ux, uy, qx, qy);
                loaded = true;
            }
            const CandRec* cr = rec + c0;
            const float2* co = kIdx ? nullptr : orig + c0;
            if (!kShared) {
                L->rec[lane] = r;
                L->orig[lane] = o;
                __builtin_amdgcn_wave_barrier();
                cr = L->rec;
                co = L->orig;
            }
            while (cm) {
                uint32_t b[4];
                bool v[4]; // wave-uniform: slot g holds a candidate
                b[0] = (uint32_t)__ffsll((unsigned long long)cm) - 1;
                cm &= cm - 1;
                v[0] = true;
#pragma unroll
                for (int g = 1; g < 4; ++g) {
                    v[g] = cm != 0;
                    b[g] = v[g] ? (uint32_t)__ffsll((unsigned long long)cm) - 1 : b[0];
                    cm &= cm - 1;
                }
                float t[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const CandRec c = cr[b[g]];
                    t[g] = fmaf(wx, c.gx, fmaf(wy, c.gy, -c.sh));
                }
                bool band[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    cnt += (v[g] && t[g] >= kCertainHi) ? 1u : 0u;
                    band[g] = v[g] && in_band(t[g]);
                }
                if (ballot(band[0] || band[1] || band[2] || band[3])) { // rare: settle in binary64
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const uint64_t bm = ballot(band[g]);
                        if (bm) {
                            float2 f;
                            if (kIdx) {
                                const uint32_t fi = __float_as_uint(cr[b[g]].pad);
                                f = make_float2(fxg[fi], fyg[fi]);
                            } else {
                                f = co[b[g]];
                            }
                            if (band[g] && d2_exact(ux, uy, f.x, f.y) < dq) ++cnt;
                            fp64 += (uint32_t)__popcll(bm & ballot(valid));
                        }
                    }
                }
            }
            if (!kShared) __builtin_amdgcn_wave_barrier();
        }
        if (ballot(valid && base + cnt < k) == 0) {
            dead = true;
            break;
        }
    }
    LeafOut out;
    out.keep = ballot(!dead && valid && base + cnt < k);
    out.fp64 = fp64;
    return out;
}

// ---------------------------------------------------------------- plain form

constexpr uint32_t kPlainRecLds = 512;

// Kept users go to out[] through ctr[0]; ids past out_cap are counted but
// not stored (the caller sees the count and reports the shortfall).
template <bool kShared>
__device__ void plain_tiles(const DsDev& ds, const QDev& q, const CandRec* __restrict__ rec,
                            const float2* __restrict__ orig, WaveLds* L, int32_t* __restrict__ out,
                            uint32_t out_cap, uint32_t* __restrict__ ctr, uint32_t& rech)
{
    const uint32_t lane = lane_id();
    const uint32_t stride = gridDim.x * kWaves;
    for (uint32_t tile = blockIdx.x * kWaves + (threadIdx.x >> 6); tile < ds.n_tiles; tile += stride) {
        const LeafOut r =
            leaf_tile<kShared, false>(ds, q.qx, q.qy, q.k, q.nc, rec, orig, tile, ds.mbr[0][tile], 0.f, 0.f, L);
        rech += r.fp64;
        if (r.keep) {
            uint32_t pos = 0;
            if (lane == 0) pos = atomicAdd(&ctr[0], (uint32_t)__popcll(r.keep));
            pos = uniform(pos);
            const uint32_t at = pos + rank_below(r.keep);
            if (((r.keep >> lane) & 1ull) && at < out_cap) out[at] = ds.uid[tile * kTile + lane];
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_plain(DsDev ds, QDev q, const CandRec* __restrict__ rec,
                                                  const float2* __restrict__ orig, int32_t* __restrict__ out,
                                                  uint32_t out_cap, uint32_t* __restrict__ ctr)
{
    __shared__ union {
        SharedRecs<kPlainRecLds> s;
        WaveLds w[kWaves];
    } lds;
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t rech = 0;
    if (q.nc <= kPlainRecLds) {
        for (uint32_t j = threadIdx.x; j < q.nc; j += kBlock) {
            lds.s.rec[j] = rec[j];
            lds.s.orig[j] = orig[j];
        }
        __syncthreads();
        plain_tiles<true>(ds, q, lds.s.rec, lds.s.orig, nullptr, out, out_cap, ctr, rech);
    } else {
        plain_tiles<false>(ds, q, rec, orig, &lds.w[wave], out, out_cap, ctr, rech);
    }
    if (lane_id() == 0 && rech) atomicAdd(&ctr[1 + (blockIdx.x * kWaves + wave) % kBuckets], rech);
}

// Candidate records of one query built on the device from facility indices
// (binary64 -> fp32 as the host does it).  A candidate that can never be
// strictly closer than q -- q itself, or a facility at q's exact position --
// gets a record that misses every user (t = -1 everywhere: never counted,
// never in the band).  The first block also clears the plain form's
// counters, so the launch that follows needs no memset.
__global__ void k_build_records(const uint32_t* __restrict__ cand, uint32_t nc, uint32_t q, float qx, float qy,
                                double wx, double wy, const float* __restrict__ fx, const float* __restrict__ fy,
                                CandRec* __restrict__ rec, float2* __restrict__ orig, uint32_t* __restrict__ ctr,
                                uint32_t n_ctr)
{
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < n_ctr; i += blockDim.x) ctr[i] = 0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nc; j += gridDim.x * blockDim.x) {
        const uint32_t f = cand[j];
        const float x = fx[f], y = fy[f];
        CandRec r;
        if (f == q || (x == qx && y == qy)) {
            r.gx = 0.f;
            r.gy = 0.f;
            r.sh = 1.f;
            r.pad = 0.f;
        } else {
            r = make_record(qx, qy, x, y, wx, wy);
        }
        rec[j] = r;
        orig[j] = make_float2(x, y);
    }
}
// ------------------------------------------------------------ count-all form
//
// Every (user, candidate) pair is evaluated: no culling, no early exit (the
// rknn-rt any-hit counter with k = infinity).  Each wave counts one group of
// U adjacent tiles and the grid has a block for every kWaves groups, so the
// hardware dispatcher balances the load: groups that cost more (near q
// every bisector crosses them and they take the banded path) hold their
// slot longer while other slots take more blocks, and the kernel's tail is
// at most one group long.  (A persistent grid with a static split left the
// waves that drew those groups running alone; work counters in global
// memory cost a serialised device atomic, about 100 ns, per group.)  The
// query's records are staged once per block in LDS (up to kCountLdsRecs of
// them; longer lists are read from global).
//
// Per 64-candidate chunk the candidates are split against the group's
// joint rectangle: "clean" ones, whose computed t is >= 19/16 (or <= -3/16)
// at the rectangle's extreme corner, give every user of the group t >= 1 (or
// <= 0), because the fp32 error is < 1/16 per evaluation; the others take
// the banded path, users on lanes, with the binary64 re-check.
//
// Counts are written in sorted user order, as 8-bit values when the query
// has fewer than 2^8 candidates and 16-bit ones below 2^16 (9 or 10 instead
// of 12 bytes of HBM traffic per user), else 32-bit.

constexpr uint32_t kCountLdsRecs = 1024;

// The candidate split of one chunk (see above): clean when certain for the
// whole rectangle (lx, ly, hx, hy relative to q).
__device__ __forceinline__ bool clean_for(const CandRec& r, float lx, float ly, float hx, float hy)
{
    const float tlo = fmaf(r.gx >= 0.f ? lx : hx, r.gx, fmaf(r.gy >= 0.f ? ly : hy, r.gy, -r.sh));
    const float thi = fmaf(r.gx >= 0.f ? hx : lx, r.gx, fmaf(r.gy >= 0.f ? hy : ly, r.gy, -r.sh));
    return tlo >= 1.1875f || thi <= -0.1875f;
}

// Banded path for n_slow candidates s[0 .. n_slow) against U tiles of users
// (lane = user), counting into acc (whole numbers, exact in fp32 below
// 2^24).  It runs in double-rate FP32 only:
//   c = clamp(4 t - 3/2)
// is exactly 1 for t >= 5/8 (certainly closer), exactly 0 for t <= 3/8
// (certainly not), and strictly between 0 and 1 inside the band, where
// 4 t - 3/2 is exact (4 t lies in [3/2, 5/2], the difference is a multiple of
// its ulp below 1).  The sum of c counts the certain cases; the sum of
// c - c^2 (each term exactly zero for c in {0, 1}, positive otherwise) is
// zero unless a band value occurred.  A tile with one re-walks its slow
// candidates with the oracle's binary64 comparison in the band.
template <int U>
__device__ __forceinline__ void count_slow(const CandRec* s, const float2* __restrict__ orig, uint32_t n_slow,
                                           const DsDev& ds,
                                           uint32_t t0, uint32_t nt, const float* wx, const float* wy, float qx,
                                           float qy, float* acc, uint32_t& rech)
{
    float sacc[U], bacc[U];
#pragma unroll
    for (int j = 0; j < U; ++j) sacc[j] = bacc[j] = 0.f;
    for (uint32_t i = 0; i < n_slow; ++i) {
        const CandRec c = s[i];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const float t = fmaf(wx[j], c.gx, fmaf(wy[j], c.gy, -c.sh));
            const float v = __builtin_amdgcn_fmed3f(fmaf(t, 4.0f, -1.5f), 0.0f, 1.0f);
            sacc[j] += v;
            bacc[j] += fmaf(v, -v, v);
        }
    }
    const uint32_t lane = lane_id();
    // one 64-bit base for the group, the tile as an immediate offset (per-tile
    // 32-bit indices would each hold a 64-bit address through the group)
    const uint64_t base = (uint64_t)t0 * kTile + lane;
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const bool band = bacc[j] > 0.f;
        if (ballot(band) == 0) {
            acc[j] += sacc[j];
            continue;
        }
        // rare: this tile met the band; recount its slow candidates exactly
        uint32_t exact = 0;
        if ((uint32_t)j < nt) {
            const float px = (ds.ux + base)[j * kTile], py = (ds.uy + base)[j * kTile];
            const double dq = d2_exact(px, py, qx, qy);
            for (uint32_t i = 0; i < n_slow; ++i) {
                const CandRec c = s[i];
                const float t = fmaf(wx[j], c.gx, fmaf(wy[j], c.gy, -c.sh));
                exact += (t >= kCertainHi) ? 1u : 0u;
                const bool inband = in_band(t);
                const uint64_t bm = ballot(inband);
                if (bm) { // the record's pad holds its candidate index
                    if (inband) {
                        const float2 f = orig[__float_as_uint(c.pad)];
                        if (d2_exact(px, py, f.x, f.y) < dq) ++exact;
                    }
                    rech += (uint32_t)__popcll(bm);
                }
            }
        }
        acc[j] += band ? (float)exact : sacc[j];
    }
}

// The users (lane = user, one per tile) and joint rectangle of group g (kept
// per group in the dataset: a union here would cost single-rate min/max and
// NaN-canonicalising instructions per group).
template <int U>
__device__ __forceinline__ void load_group(const DsDev& ds, uint32_t g, float* px, float* py, float4& m)
{
    static_assert(U == 4 || U == 6 || U == 8, "group rectangles are kept for 4-, 6- and 8-tile groups");
    const uint32_t lane = lane_id();
    const uint32_t t0 = g * U;
    const uint32_t nt = uniform(min((uint32_t)U, ds.n_tiles - t0));
    m = (U == 4 ? ds.gmbr4 : U == 6 ? ds.gmbr6 : ds.gmbr8)[g];
    const uint64_t base = (uint64_t)t0 * kTile + lane;
    const float* ux = ds.ux + base;
    const float* uy = ds.uy + base;
    if (nt == (uint32_t)U) { // a full group: one address, immediate offsets
#pragma unroll
        for (int j = 0; j < U; ++j) {
            px[j] = ux[j * kTile];
            py[j] = uy[j * kTile];
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        px[j] = py[j] = 0.f;
        if ((uint32_t)j < nt) {
            px[j] = ux[j * kTile];
            py[j] = uy[j * kTile];
        }
    }
}

template <int U, bool kLdsRecs, typename OutT>
__device__ __forceinline__ void count_group(const DsDev& ds, const QDev& q, uint32_t nc,
                                            const CandRec* __restrict__ srec, const CandRec* __restrict__ rec,
                                            const float2* __restrict__ orig, OutT* __restrict__ counts,
                                            CandRec* wl, uint32_t t0, uint32_t nt, const float* px,
                                            const float* py, float4 m, uint32_t& rech)
{
    const uint32_t lane = lane_id();
    float wx[U], wy[U], acc[U]; // acc: the count so far, a whole number
#pragma unroll
    for (int j = 0; j < U; ++j) {
        wx[j] = px[j] - q.qx;
        wy[j] = py[j] - q.qy;
        acc[j] = 0.f;
    }
    // wave-uniform: kept in scalar registers through the chunks
    const float lx = uniformf(m.x - q.qx), ly = uniformf(m.y - q.qy), hx = uniformf(m.z - q.qx),
                hy = uniformf(m.w - q.qy);
    for (uint32_t c0 = 0; c0 < nc; c0 += 64) {
        const uint32_t jc = c0 + lane;
        const bool have = jc < nc;
        // the records' place is a template parameter: a loop holding both
        // paths was compiled with a wait for every outstanding load
        // (vmcnt(0) lgkmcnt(0)) at each chunk, the global path's loads
        // sharing the LDS path's registers
        CandRec r = {0.f, 0.f, 0.f, 0.f};
        if (have) r = kLdsRecs ? srec[jc] : rec[jc];
        const bool clean = have && clean_for(r, lx, ly, hx, hy);
        const bool slow = have && !clean;
        const uint64_t mc = ballot(clean), ms = ballot(slow);
        if (clean) wl[rank_below(mc)] = r;
        if (slow) { // the facility's coordinates are read only if a band value occurs
            r.pad = __uint_as_float(jc);
            wl[64 + rank_below(ms)] = r;
        }
        __builtin_amdgcn_wave_barrier();
        const uint32_t n_clean = (uint32_t)__popcll(mc), n_slow = (uint32_t)__popcll(ms);
#pragma unroll 4
        for (uint32_t i = 0; i < n_clean; ++i) {
            const CandRec c = wl[i];
#pragma unroll
            for (int j = 0; j < U; ++j) // med3(t,0,1) folds into the FMA's clamp bit
                acc[j] += __builtin_amdgcn_fmed3f(fmaf(wx[j], c.gx, fmaf(wy[j], c.gy, -c.sh)), 0.0f, 1.0f);
        }
        if (n_slow) count_slow<U>(wl + 64, orig, n_slow, ds, t0, nt, wx, wy, q.qx, q.qy, acc, rech);
        __builtin_amdgcn_wave_barrier();
    }
    // the padding users of the last tile are not stored (the output may be
    // the caller's array of exactly nu counts)
    OutT* __restrict__ out = counts + ((uint64_t)t0 * kTile + lane);
    const uint32_t u0 = t0 * kTile + lane;
#pragma unroll
    for (int j = 0; j < U; ++j)
        if ((uint32_t)j < nt && u0 + (uint32_t)j * kTile < ds.nu) out[j * kTile] = (OutT)(uint32_t)acc[j];
}

// The first group's users are requested before the records are staged, so
// the two loads overlap.  (Waves that request their next group's users while
// counting the current one measured slower: over a persistent grid, the
// strided walk leaves a longer tail than the hardware dispatcher; over runs
// of 2 or 4 adjacent groups per wave, 41.7 and 35.1 TFLOP/s against 45.8 for
// one group per wave at k = 10.)  W is the occupancy the register
// allocation may be held to (W = 1 leaves it free: 53, 60 and 90 VGPRs for
// U = 4, 6 and 8).
template <int U, int W, typename OutT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W)))
void k_count_all(DsDev ds, QDev q, const CandRec* __restrict__ rec, const float2* __restrict__ orig,
                 OutT* __restrict__ counts, uint32_t* __restrict__ rech_out)
{
    extern __shared__ CandRec s_recs[];       // the query's records (nc <= kCountLdsRecs)
    __shared__ CandRec s_w[kWaves][128];      // per chunk: clean ones at [0, 64), the others at [64, 128)
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nc = uniform(q.nc);
    const bool lds_recs = nc <= kCountLdsRecs;
    const uint32_t n_groups = (ds.n_tiles + U - 1) / U;
    const uint32_t stride = gridDim.x * kWaves;
    uint32_t g = uniform(blockIdx.x * kWaves + wave);
    float px[U], py[U];
    float4 m = make_float4(0.f, 0.f, 0.f, 0.f);
    if (g < n_groups) load_group<U>(ds, g, px, py, m);
    if (lds_recs) {
        for (uint32_t i = threadIdx.x; i < nc; i += kBlock) s_recs[i] = rec[i];
        __syncthreads();
    }
    uint32_t rech = 0;
    while (g < n_groups) {
        const uint32_t t0 = g * U, nt = min((uint32_t)U, ds.n_tiles - t0);
        if (lds_recs)
            count_group<U, true, OutT>(ds, q, nc, s_recs, rec, orig, counts, s_w[wave], t0, nt, px, py, m, rech);
        else
            count_group<U, false, OutT>(ds, q, nc, s_recs, rec, orig, counts, s_w[wave], t0, nt, px, py, m, rech);
        g += stride;
        if (g < n_groups) load_group<U>(ds, g, px, py, m);
    }
    if (lane_id() == 0 && rech) atomicAdd(&rech_out[(blockIdx.x * kWaves + wave) % kBuckets], rech);
}

// Count-all over a set of queries in one launch.  Each wave loads its
// group's users once and counts them against every query of the set in
// turn; the block stages all the set's records (at most kCountSetRecs) in
// LDS once.  A wave's start (argument and user loads, record staging) and
// the launch's first and last rounds are then paid once per set instead of
// once per query.  Query i's counts go to counts + qs[i].coff, its records
// are rec[qs[i].roff ..) and orig[qs[i].roff ..).
template <int U, typename OutT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(U == 4 ? 8 : U == 6 ? 6 : 5)))
void k_count_set(DsDev ds, const CountSetQ* __restrict__ qs, uint32_t nq, uint32_t n_recs,
                 const CandRec* __restrict__ rec, const float2* __restrict__ orig, OutT* __restrict__ counts,
                 uint32_t* __restrict__ rech_out)
{
    extern __shared__ CandRec s_recs[];   // every query's records, query-major
    __shared__ CandRec s_w[kWaves][128];  // per chunk: clean ones at [0, 64), the others at [64, 128)
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t n_groups = (ds.n_tiles + U - 1) / U;
    const uint32_t stride = gridDim.x * kWaves;
    uint32_t g = uniform(blockIdx.x * kWaves + wave);
    float px[U], py[U];
    float4 m = make_float4(0.f, 0.f, 0.f, 0.f);
    if (g < n_groups) load_group<U>(ds, g, px, py, m);
    for (uint32_t i = threadIdx.x; i < n_recs; i += kBlock) s_recs[i] = rec[i];
    __syncthreads();
    uint32_t rech = 0;
    while (g < n_groups) {
        const uint32_t t0 = g * U, nt = min((uint32_t)U, ds.n_tiles - t0);
        for (uint32_t i = 0; i < nq; ++i) {
            const CountSetQ d = qs[i]; // uniform: scalar loads
            QDev q{};
            q.qx = d.qx;
            q.qy = d.qy;
            q.nc = d.nc;
            count_group<U, true, OutT>(ds, q, d.nc, s_recs + d.roff, rec + d.roff, orig + d.roff,
                                       counts + d.coff, s_w[wave], t0, nt, px, py, m, rech);
        }
        g += stride;
        if (g < n_groups) load_group<U>(ds, g, px, py, m);
    }
    if (lane_id() == 0 && rech) atomicAdd(&rech_out[(blockIdx.x * kWaves + wave) % kBuckets], rech);
}

template <typename SrcT>
__global__ void k_scatter_u32(const SrcT* __restrict__ src, const uint32_t* __restrict__ perm, uint32_t n,
                              uint32_t* __restrict__ dst)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[perm[i]] = (uint32_t)src[i];
}

// ------------------------------------------------------ dataset construction

// Hilbert index of a 16-bit x 16-bit cell (the iterative quadrant walk of
// Hilbert's curve).  Consecutive keys are adjacent cells, so 64 consecutive
// users form a compact tile; a Z-order (Morton) key jumps between quadrants
// and gives some tiles rectangles as wide as the whole map.
__device__ __forceinline__ uint32_t hilbert16(uint32_t x, uint32_t y)
{
    uint32_t d = 0;
    for (uint32_t s = 1u << 15; s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u;
        const uint32_t ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) { // rotate the quadrant
            if (rx == 1) {
                x = 0xffffu - x;
                y = 0xffffu - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

__global__ void k_curve_keys(const float* __restrict__ ux, const float* __restrict__ uy, uint32_t n, float minx,
                             float miny, float sx, float sy, uint32_t* __restrict__ keys, uint32_t* __restrict__ idx)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float fx = (ux[i] - minx) * sx, fy = (uy[i] - miny) * sy;
        const uint32_t ix = (uint32_t)fminf(fmaxf(fx, 0.f), 65535.f);
        const uint32_t iy = (uint32_t)fminf(fmaxf(fy, 0.f), 65535.f);
        keys[i] = hilbert16(ix, iy);
        idx[i] = i;
    }
}

__global__ void k_gather_users(const uint32_t* __restrict__ order, uint32_t n, uint32_t n_pad,
                               const float* __restrict__ ux_in, const float* __restrict__ uy_in,
                               const int32_t* __restrict__ uid_in, float* __restrict__ ux, float* __restrict__ uy,
                               int32_t* __restrict__ uid)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += gridDim.x * blockDim.x) {
        const uint32_t s = order[i < n ? i : n - 1];
        ux[i] = ux_in[s];
        uy[i] = uy_in[s];
        uid[i] = i < n ? uid_in[s] : -1;
    }
}

__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__global__ __launch_bounds__(kBlock) void k_tile_mbr(const float* __restrict__ ux, const float* __restrict__ uy,
                                                     uint32_t n_tiles, float4* __restrict__ mbr)
{
    const uint32_t lane = lane_id();
    for (uint32_t t = blockIdx.x * kWaves + (threadIdx.x >> 6); t < n_tiles; t += gridDim.x * kWaves) {
        const float x = ux[t * kTile + lane], y = uy[t * kTile + lane];
        const float a = wave_min(x), b = wave_min(y), c = wave_max(x), d = wave_max(y);
        if (lane == 0) mbr[t] = make_float4(a, b, c, d);
    }
}

__global__ __launch_bounds__(kBlock) void k_level_mbr(const float4* __restrict__ child, uint32_t n_child,
                                                      float4* __restrict__ parent, uint32_t n_parent)
{
    const uint32_t lane = lane_id();
    for (uint32_t p = blockIdx.x * kWaves + (threadIdx.x >> 6); p < n_parent; p += gridDim.x * kWaves) {
        uint32_t c = p * kFanout + lane;
        // lanes past the last child repeat the node's first child: no effect on min/max
        const float4 m = child[c < n_child ? c : p * kFanout];
        const float a = wave_min(m.x), b = wave_min(m.y), cc = wave_max(m.z), d = wave_max(m.w);
        if (lane == 0) parent[p] = make_float4(a, b, cc, d);
    }
}

__global__ void k_group_mbr(const float4* __restrict__ tiles, uint32_t n_tiles, uint32_t fan,
                            float4* __restrict__ groups)
{
    const uint32_t n_groups = (n_tiles + fan - 1) / fan;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n_groups; g += gridDim.x * blockDim.x) {
        float4 m = tiles[g * fan];
        const uint32_t e = min(n_tiles, (g + 1) * fan);
        for (uint32_t t = g * fan + 1; t < e; ++t) {
            const float4 b = tiles[t];
            m.x = fminf(m.x, b.x);
            m.y = fminf(m.y, b.y);
            m.z = fmaxf(m.z, b.z);
            m.w = fmaxf(m.w, b.w);
        }
        groups[g] = m;
    }
}

// -------------------------------------------------------------- batched form
//
// One block of kW (8 or 4) waves per query.  The block builds the query's
// candidate records, walks the bounding-rectangle tree breadth-first
// through per-level frontiers in LDS, tests the surviving tiles and appends
// kept users to the query's result (in LDS, spilling to its segment), which
// it then sorts and writes to pinned host memory with the query's words.
// Every counter is an LDS counter of the block; the only global atomics
// are the landing reservation (one per query) and the block's completion
// ticket, so queries proceed independently.
//
// Frontiers are bounded: the deepest level with pending nodes is always
// expanded first, at most kPhaseItems nodes at a time, so a level's frontier
// is empty when its parent level is expanded and never holds more than
// kPhaseItems * 64 = kFront nodes.  A phase with fewer nodes than waves
// gives each node several waves, each counting a slice of the candidates;
// the slices' counts meet in LDS (cover[]).  Without the split a phase of
// one node (the root, say) runs on one wave of the eight while the others
// wait at the barrier.
//
// Shared records are sorted by the signs of (gx, gy), so that within a run
// the corner of a rectangle minimising t is the same for every candidate:
// the node test is then two FMAs, a compare and an add per (child,
// candidate).

constexpr uint32_t kQWaves = kQueryBlock / 64; // waves of the largest block shape
constexpr uint32_t kPhaseItems = kFront / kFanout;
constexpr uint32_t kRecLds = 1024; // shared-record capacity of the batched form

struct QueryCtl {
    uint32_t n[kMaxLevels];     // frontier fill per level
    uint32_t c[kMaxLevels];     // frontier cursor per level
    uint32_t items[kMaxLevels]; // nodes that entered each level's frontier
    uint32_t nres;              // result ids of the query so far
    uint32_t fp64;              // (user, candidate) tests settled in binary64
    uint32_t qcnt[4];           // shared records per sign quadrant
    uint32_t qcur[4];           // placement cursors per quadrant
    uint32_t quad[5];           // quadrant q holds records [quad[q], quad[q+1])
    uint32_t red[2 * kQWaves + 1]; // the segment sort's reductions and skew flag
    uint32_t land;              // the query's landing offset (block broadcast)
    uint32_t pad[1];
};
constexpr uint32_t kCtlWords = sizeof(QueryCtl) / 4;
constexpr size_t kCtlBytes = (sizeof(QueryCtl) + 15) / 16 * 16;
constexpr size_t kCoverBytes = kPhaseItems * 64 * sizeof(uint32_t);
// candidate area: the shared records, or one streaming slot per wave; after
// the descent it holds the in-block sort of the result ids
constexpr size_t kCandBytes = sizeof(SharedRecsIdx<kRecLds>) > kQWaves * sizeof(WaveLds)
                                  ? sizeof(SharedRecsIdx<kRecLds>)
                                  : kQWaves * sizeof(WaveLds);
constexpr size_t kResBytes = kResLds * sizeof(uint32_t); // the query's result ids
// After the descent the block bucket-sorts up to 8 result ids per thread in
// registers and LDS (from the cover counts up to the frontiers, free by
// then); larger results, or ids bunched in one bucket, are radix-sorted
// through the query's segment.
constexpr int kInItems = 8;
constexpr uint32_t kResBuckets = 1024; // buckets of the in-block sort
static_assert((kQueryBlock * kInItems + kResBuckets) * sizeof(uint32_t) <= kCoverBytes + kCandBytes + kResBytes,
              "in-block sort storage");

// Sign quadrant of a record: bit 0 gx < 0, bit 1 gy < 0 (the complement of
// the corner choice gx >= 0 ? low : high of the node test).
__device__ __forceinline__ uint32_t quadrant(const CandRec& r)
{
    return (r.gx >= 0.f ? 0u : 1u) | (r.gy >= 0.f ? 0u : 2u);
}

// Per lane (child), count candidates [i0, i1) of one quadrant run whose t at
// corner (cx, cy) is >= 5/8.  Stops once every valid lane has >= k.
__device__ __forceinline__ void cover_run(const CandRec* __restrict__ cr, uint32_t i0, uint32_t i1, float cx,
                                          float cy, uint32_t k, bool valid, uint32_t& full, bool& live)
{
    uint32_t i = i0;
    while (i < i1 && live) {
        if (i + 16 <= i1) { // unrolled so the LDS reads pipeline
#pragma unroll
            for (uint32_t d = 0; d < 16; ++d) {
                const CandRec r = cr[i + d];
                full += (fmaf(cx, r.gx, fmaf(cy, r.gy, -r.sh)) >= kCertainHi) ? 1u : 0u;
            }
            i += 16;
        } else {
            for (; i < i1; ++i) {
                const CandRec r = cr[i];
                full += (fmaf(cx, r.gx, fmaf(cy, r.gy, -r.sh)) >= kCertainHi) ? 1u : 0u;
            }
        }
        live = ballot(valid && full < k) != 0;
    }
}

// Lanes = the <= 64 children of `node` at a level with MBRs `cm`: per child,
// how many of candidates [lo, hi) certainly cover its whole rectangle (the
// internal-node test of rknn-inf/inf.h:219-237, with certified corner
// evaluation instead of a 4-corner binary64 test).  A count that reaches k
// may stop early: the child is culled whatever the other candidates say.
//   kShared: quadrant-sorted records in LDS (quad = run starts)
//   else:    records in global memory, streamed through wlds
template <bool kShared>
__device__ uint32_t cover_count(const float4* __restrict__ cm, uint32_t n_child, uint32_t node, const QDev& Q,
                                const CandRec* __restrict__ rec, const uint32_t* quad, uint32_t lo, uint32_t hi,
                                CandRec* wlds)
{
    const uint32_t lane = lane_id();
    const uint32_t c = node * kFanout + lane;
    const bool valid = c < n_child;
    const float4 m = cm[valid ? c : 0];
    const float lx = m.x - Q.qx, ly = m.y - Q.qy, hx = m.z - Q.qx, hy = m.w - Q.qy;
    uint32_t full = 0;
    bool live = true;
    if (kShared) {
        for (uint32_t qd = 0; qd < 4 && live; ++qd) {
            const uint32_t a = max(lo, uniform(quad[qd])), e = min(hi, uniform(quad[qd + 1]));
            if (a < e) cover_run(rec, a, e, (qd & 1) ? hx : lx, (qd & 2) ? hy : ly, Q.k, valid, full, live);
        }
        return full;
    }
    for (uint32_t c0 = lo; c0 < hi && live; c0 += 64) {
        const uint32_t n = min(64u, hi - c0);
        __builtin_amdgcn_wave_barrier();
        if (lane < n) wlds[lane] = rec[c0 + lane];
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i0 = 0; i0 < n && live; i0 += 16) {
            const uint32_t i1 = min(n, i0 + 16);
            for (uint32_t i = i0; i < i1; ++i) {
                const CandRec r = wlds[i];
                const float t = fmaf(r.gx >= 0.f ? lx : hx, r.gx, fmaf(r.gy >= 0.f ? ly : hy, r.gy, -r.sh));
                full += (t >= kCertainHi) ? 1u : 0u;
            }
            live = ballot(valid && full < Q.k) != 0;
        }
    }
    return full;
}

// Append the surviving children of `node` to a frontier (one LDS atomic per wave).
__device__ __forceinline__ void push_children(uint64_t surv, uint32_t node, uint32_t* front, uint32_t* fill)
{
    const uint32_t ns = (uint32_t)__popcll(surv);
    if (!ns) return;
    const uint32_t lane = lane_id();
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(fill, ns);
    pos = uniform(pos);
    if ((surv >> lane) & 1ull) front[pos + rank_below(surv)] = node * kFanout + lane;
}

// Frontier of level l in the block's LDS: kFront entries per level below the
// top, 64 for the top level.
__device__ __forceinline__ uint32_t* frontier(uint32_t* front, uint32_t l) { return front + l * kFront; }

// The descent of one query (all threads of the block call it).  The first
// phase expands a virtual root whose children are the top-level nodes.
// Kept users go to the block's LDS result buffer res[] (the first kResLds of
// them) and past it to the query's segment in HBM at the same positions.
// Shared records (kShared) carry their facility index for the band tests.
template <uint32_t kW, bool kShared>
__device__ void descend(const DsDev& ds, const BatchDev& b, const QDev& Q, uint32_t qi, QueryCtl& ctl,
                        uint32_t* front, uint32_t* cover, const CandRec* __restrict__ rec,
                        const float2* __restrict__ orig, WaveLds* L, uint32_t* res)
{
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = lane_id();
    const uint32_t top = uniform(ds.n_levels - 1);
    int32_t* __restrict__ seg = b.seg + (uint64_t)qi * b.seg_cap;
    uint32_t fp64 = 0; // per wave
    bool root = true;
    for (;;) {
        uint32_t l = top + 1, c = 0, n = 1; // the root phase
        if (!root) {
            l = kMaxLevels; // deepest level with pending nodes
            for (uint32_t p = 0; p <= top; ++p)
                if (ctl.c[p] < ctl.n[p]) {
                    l = p;
                    break;
                }
            l = uniform(l);
            if (l == kMaxLevels) break;
            c = uniform(ctl.c[l]);
            n = uniform(ctl.n[l]);
        }
        const uint32_t* fl = frontier(front, l);
        uint32_t end = n;
        if (l == 0) {
            // every pending tile; the next tile's rectangle and users are
            // loaded while the current one is counted
            uint32_t i = c + wave;
            uint32_t tile = 0;
            float4 m = make_float4(0.f, 0.f, 0.f, 0.f);
            float ux = 0.f, uy = 0.f;
            if (i < n) {
                tile = uniform(fl[i]);
                m = ds.mbr[0][tile];
                ux = ds.ux[tile * kTile + lane];
                uy = ds.uy[tile * kTile + lane];
            }
            for (; i < n; i += kW) {
                uint32_t tn = 0;
                float4 mn = make_float4(0.f, 0.f, 0.f, 0.f);
                float uxn = 0.f, uyn = 0.f;
                if (i + kW < n) {
                    tn = uniform(fl[i + kW]);
                    mn = ds.mbr[0][tn];
                    uxn = ds.ux[tn * kTile + lane];
                    uyn = ds.uy[tn * kTile + lane];
                }
                const LeafOut r = leaf_tile<kShared, true, kShared>(ds, Q.qx, Q.qy, Q.k, Q.nc, rec, orig, tile, m, ux,
                                                                    uy, L, b.fx, b.fy);
                fp64 += r.fp64;
                if (r.keep) {
                    uint32_t pos = 0;
                    if (lane == 0) pos = atomicAdd(&ctl.nres, (uint32_t)__popcll(r.keep));
                    pos = uniform(pos);
                    if ((r.keep >> lane) & 1ull) {
                        const uint32_t rk = pos + rank_below(r.keep);
                        const int32_t id = ds.uid[tile * kTile + lane];
                        if (rk < kResLds)
                            res[rk] = (uint32_t)id;
                        else if (rk < b.seg_cap)
                            seg[rk] = id;
                    }
                }
                tile = tn;
                m = mn;
                ux = uxn;
                uy = uyn;
            }
        } else {
            // nodes [c, end) of level l; their children (level l-1) fill
            // that level's frontier; S waves per node, one candidate slice each
            end = min(n, c + kPhaseItems);
            const uint32_t items = end - c;
            const uint32_t S = max(1u, kW / items);
            const uint32_t per = ((Q.nc + S - 1) / S + 15) & ~15u;
            const uint32_t cl = l - 1;
            const float4* cmbr = ds.mbr[cl];
            const uint32_t n_child = ds.nlev[cl];
            uint32_t* dst = frontier(front, cl);
            for (uint32_t w = wave; w < items * S; w += kW) {
                const uint32_t it = w / S, sl = w % S;
                const uint32_t node = root ? 0u : uniform(fl[c + it]);
                const uint32_t lo = min(Q.nc, sl * per), hi = min(Q.nc, lo + per);
                const uint32_t full = cover_count<kShared>(cmbr, n_child, node, Q, rec, ctl.quad, lo, hi, L->rec);
                if (S == 1)
                    push_children(ballot(node * kFanout + lane < n_child && full < Q.k), node, dst, &ctl.n[cl]);
                else if (lo < hi)
                    atomicAdd(&cover[it * 64 + lane], full);
            }
            if (S > 1) {
                __syncthreads();
                for (uint32_t it = wave; it < items; it += kW) {
                    const uint32_t node = root ? 0u : uniform(fl[c + it]);
                    const uint32_t full = cover[it * 64 + lane];
                    cover[it * 64 + lane] = 0;
                    push_children(ballot(node * kFanout + lane < n_child && full < Q.k), node, dst, &ctl.n[cl]);
                }
            }
        }
        __syncthreads();
        if (tid == 0 && !root) {
            if (end == n) { // frontier drained
                ctl.items[l] += n;
                ctl.c[l] = 0;
                ctl.n[l] = 0;
            } else {
                ctl.c[l] = end;
            }
        }
        __syncthreads();
        root = false;
    }
    if (lane == 0 && fp64) atomicAdd(&ctl.fp64, fp64);
}

// ------------------------------------------------------ result-segment sort
//
// A query's result ids are distinct, so an unstable bucket sort is exact:
// keys are spread over kNB buckets by their position in [min, max] (a
// monotone map, so bucket order is key order), placed with one LDS atomic
// each, and every bucket (about n / kNB keys) is finished by an insertion
// sort.  Cost O(n) with six block barriers, against six 4-bit passes of a
// block radix sort.  A bucket holding more than kMaxBucket keys (ids bunched
// very unevenly inside their range) aborts the attempt; the segment is then
// left untouched for the radix-sort fallback.
constexpr uint32_t kMaxBucket = 32;

// Wave-wide prefix operations on DPP: row shifts 1, 2, 4, 8 within each row
// of 16 lanes, then row 15 broadcast into rows 1 and 3 and row 31 into rows
// 2 and 3.  No lane-address registers are needed (a shuffle network keeps one
// per step live, which spilled in the batched traversal kernel).  A lane
// reading past its row, or in a row the mask leaves out, takes `id`.
constexpr int kDppShr1 = 0x111, kDppShr2 = 0x112, kDppShr4 = 0x114, kDppShr8 = 0x118;
constexpr int kDppBcast15 = 0x142, kDppBcast31 = 0x143;

template <typename Op>
__device__ __forceinline__ uint32_t wave_prefix(uint32_t x, uint32_t id, Op op)
{
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, kDppShr1, 0xf, 0xf, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, kDppShr2, 0xf, 0xf, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, kDppShr4, 0xf, 0xf, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, kDppShr8, 0xf, 0xf, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, kDppBcast15, 0xa, 0xf, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, kDppBcast31, 0xc, 0xf, false));
    return x; // lane i: op over lanes 0..i
}

// inclusive prefix sum over the wave
__device__ __forceinline__ uint32_t wave_scan_incl_u32(uint32_t v)
{
    return wave_prefix(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
    const uint32_t p = wave_prefix(v, 0xffffffffu, [](uint32_t a, uint32_t b) { return min(a, b); });
    return (uint32_t)__builtin_amdgcn_readlane((int)p, 63);
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
    const uint32_t p = wave_prefix(v, 0u, [](uint32_t a, uint32_t b) { return max(a, b); });
    return (uint32_t)__builtin_amdgcn_readlane((int)p, 63);
}

// Sorts n distinct keys (unsigned order) held in registers -- key[i] of
// thread t is item t + i * kThreads; items >= n are ignored -- into out[0, n)
// in LDS.  LDS: out[n], cnt[kNB], red[2 * kThreads / 64 + 1].  All threads
// call it.  Returns false, out[] then undefined, when a bucket held more than
// kMaxBucket keys (ids bunched very unevenly in their range).
template <int kThreads, int kItems, uint32_t kNB>
__device__ bool bucket_sort_regs(const uint32_t (&key)[kItems], uint32_t n, uint32_t* out, uint32_t* cnt,
                                 uint32_t* red)
{
    constexpr uint32_t kW = kThreads / 64;
    constexpr uint32_t kPer = kNB / kThreads;
    static_assert(kPer * kThreads == kNB, "buckets per thread");
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = lane_id();
    uint32_t pos[kItems];
    uint32_t lo = 0xffffffffu, hi = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        pos[i] = 0;
        if (tid + i * kThreads < n) {
            lo = min(lo, key[i]);
            hi = max(hi, key[i]);
        }
    }
    for (uint32_t b = tid; b < kNB; b += kThreads) cnt[b] = 0;
    lo = wave_min_u32(lo);
    hi = wave_max_u32(hi);
    if (lane == 0) {
        red[wave] = lo;
        red[kW + wave] = hi;
    }
    if (tid == 0) red[2 * kW] = 0;
    __syncthreads();
    lo = red[0];
    hi = red[kW];
    for (uint32_t w = 1; w < kW; ++w) {
        lo = min(lo, red[w]);
        hi = max(hi, red[kW + w]);
    }
    const float scale = (float)kNB / ((float)(hi - lo) + 1.0f);
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        if (tid + i * kThreads < n) {
            const uint32_t b = min(kNB - 1, (uint32_t)((float)(key[i] - lo) * scale));
            pos[i] = (b << 16) | atomicAdd(&cnt[b], 1u);
        }
    }
    __syncthreads();
    // exclusive scan of the counts: kPer consecutive buckets per thread
    uint32_t local[kPer], sum = 0;
#pragma unroll
    for (uint32_t p = 0; p < kPer; ++p) {
        local[p] = cnt[tid * kPer + p];
        sum += local[p];
    }
    const uint32_t inc = wave_scan_incl_u32(sum);
    if (lane == 63) red[wave] = inc;
    __syncthreads();
    uint32_t run = inc - sum;
    for (uint32_t w = 0; w < wave; ++w) run += red[w];
#pragma unroll
    for (uint32_t p = 0; p < kPer; ++p) {
        cnt[tid * kPer + p] = run; // bucket starts
        run += local[p];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kItems; ++i)
        if (tid + i * kThreads < n) out[cnt[pos[i] >> 16] + (pos[i] & 0xffffu)] = key[i];
    __syncthreads();
    for (uint32_t b = tid; b < kNB; b += kThreads) {
        const uint32_t s0 = cnt[b], e0 = b + 1 < kNB ? cnt[b + 1] : n;
        if (e0 - s0 > kMaxBucket) {
            red[2 * kW] = 1;
            continue;
        }
        for (uint32_t x = s0 + 1; x < e0; ++x) {
            const uint32_t v = out[x];
            uint32_t y = x;
            while (y > s0 && out[y - 1] > v) {
                out[y] = out[y - 1];
                --y;
            }
            out[y] = v;
        }
    }
    __syncthreads();
    return red[2 * kW] == 0;
}

// v itself, opaque to the compiler: addresses formed from it inside a loop
// are not hoisted out of the loop (where, kept for the whole kernel, they
// spilled to scratch: 56 bytes per lane, 29 MB written per 1024-query
// dispatch)
__device__ __forceinline__ uint32_t opaque_u32(uint32_t v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Takes n ids of the landing area for one query: the offset, or kNotLanded
// when the area cannot hold them (the run is then repeated with a larger
// area).
__device__ __forceinline__ uint32_t reserve_landing(const BatchDev& b, uint32_t n)
{
    const unsigned long long s = atomicAdd(&b.land_ctr[0], (unsigned long long)n);
    return s + n <= b.land_cap ? (uint32_t)s : kNotLanded;
}

// Query qi's four words in host memory: result count, landing offset,
// candidates (the device filter's count when it made the list: beyond the
// list's segment when it overflowed), flags.
__device__ __forceinline__ void put_words(const BatchDev& b, uint32_t qi, uint32_t n, uint32_t start, const QDev& Q)
{
    b.hqres[qi] = n;
    b.hqres[b.nq + qi] = start;
    b.hqres[2 * b.nq + qi] = b.dn ? b.dn[qi] : Q.nc;
    b.hqres[3 * b.nq + qi] = Q.trivial ? kQflagTrivial : 0u;
}

// Stable LSD radix sort for results the block's in-LDS sort does not take
// (more ids than it holds, or ids bunched in one of its buckets): 8-bit
// digits over the bits below the largest id, passes between the query's
// segment and its range of the device landing area.  Each of the block's kW
// waves owns a contiguous kW-th of the keys: it counts its digits, and the
// block turns the counts into a start per (wave, digit) (digit order first,
// wave order within a digit).  Then each wave scatters its keys 64 at a
// time in order, a key's place among the equal digits of its chunk coming
// from eight ballots.  The cost does not depend on how the ids are spread.
constexpr uint32_t kRadixDigits = 256;

template <uint32_t kW>
__device__ void radix_pass(const int32_t* __restrict__ src, int32_t* __restrict__ dst, uint32_t n, uint32_t per,
                           uint32_t shift, uint32_t flip, uint32_t (*hist)[kRadixDigits], uint32_t* red,
                           uint32_t tid, uint32_t wave, uint32_t lane)
{
    static_assert(kW * 64 >= kRadixDigits, "a thread per digit");
    const uint32_t lo = min(n, wave * per), hi = min(n, lo + per);
    for (uint32_t d = lane; d < kRadixDigits; d += 64) hist[wave][d] = 0; // the wave's own row
    for (uint32_t j = lo + lane; j < hi; j += 64) atomicAdd(&hist[wave][(((uint32_t)src[j] ^ flip) >> shift) & 255u], 1u);
    __syncthreads();
    // thread d < 256 (the first four waves): digit d's total, its start (a
    // scan over digits) and each wave's start within the digit (a scan over
    // waves)
    uint32_t tot = 0, inc = 0;
    if (tid < kRadixDigits) {
        for (uint32_t w = 0; w < kW; ++w) {
            const uint32_t c = hist[w][tid];
            hist[w][tid] = tot;
            tot += c;
        }
        inc = tot;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = (uint32_t)__shfl_up((int)inc, o);
            if ((int)lane >= o) inc += v;
        }
        if (lane == 63) red[wave] = inc;
    }
    __syncthreads();
    if (tid < kRadixDigits) {
        uint32_t base = inc - tot;
        for (uint32_t w = 0; w < wave; ++w) base += red[w];
        for (uint32_t w = 0; w < kW; ++w) hist[w][tid] += base;
    }
    __syncthreads();
    for (uint32_t c0 = lo; c0 < hi; c0 += 64) {
        const uint32_t j = c0 + lane;
        const bool valid = j < hi;
        const int32_t key = valid ? src[j] : 0;
        const uint32_t d = (((uint32_t)key ^ flip) >> shift) & 255u;
        uint64_t peers = ballot(valid); // lanes of this chunk with the same digit
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const uint64_t m = ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? m : ~m;
        }
        const uint32_t r = rank_below(peers);
        if (valid) dst[hist[wave][d] + r] = key;
        if (valid && (peers >> lane) == 1) hist[wave][d] += r + 1; // the digit's last lane moves its start
    }
    __syncthreads();
}

// Sort the n ids of a query's segment into its landing range [start,
// start + n) on the device, then write them to the same range of the host
// area (16-byte lanes from the first 16-byte boundary; both areas are
// aligned alike).  Ids are below 2^id_bits (non-negative), or any int32
// (id_bits = 32: the sign bit flipped so that unsigned digit order is id
// order); the passes alternate so that the last one ends in the landing
// range.  Called by the whole block, with the thread's index, wave and lane.
template <uint32_t kW>
__device__ void radix_land(const BatchDev& b, int32_t* seg, uint32_t n, uint32_t start,
                           uint32_t (*hist)[kRadixDigits], uint32_t* red, uint32_t tid, uint32_t wave, uint32_t lane)
{
    constexpr uint32_t kThreads = kW * 64;
    const uint32_t bits = b.id_bits < 32 ? b.id_bits : 32;
    const uint32_t flip = bits == 32 ? 0x80000000u : 0u;
    const uint32_t passes = (bits + 7) / 8;
    const uint32_t per = (n + kW - 1) / kW;
    int32_t* dst = b.dland + start;
    int32_t* from = seg;
    int32_t* to = dst;
    if ((passes & 1u) == 0) { // an even pass count ends where it starts: start in the landing range
        for (uint32_t i = tid; i < n; i += kThreads) dst[i] = seg[i];
        __syncthreads();
        from = dst;
        to = seg;
    }
    for (uint32_t p = 0; p < passes; ++p) {
        radix_pass<kW>(from, to, n, per, 8 * p, flip, hist, red, tid, wave, lane);
        int32_t* t = from;
        from = to;
        to = t;
    }
    int32_t* host = b.hland + start;
    const uint32_t head = min(n, (4u - (start & 3u)) & 3u);
    if (tid < head) host[tid] = dst[tid];
    const uint32_t nv = (n - head) / 4;
    const int4* s4 = reinterpret_cast<const int4*>(dst + head);
    int4* d4 = reinterpret_cast<int4*>(host + head);
    for (uint32_t i = tid; i < nv; i += kThreads) d4[i] = s4[i];
    for (uint32_t t = head + nv * 4 + tid; t < n; t += kThreads) host[t] = dst[t];
}

// kW waves per query.  With 8-wave blocks at eight waves per SIMD (<= 64
// VGPRs), a batch of 1024 queries is resident at once on 256 CUs; 4-wave
// blocks may use twice the registers.  The autotuner picks the shape per
// workload.  A result of up to kResLds ids never touches HBM: it is
// collected in LDS, sorted there and written once, coalesced, to the
// landing area in host memory.
template <uint32_t kW>
__global__ __launch_bounds__(kW * 64) __attribute__((amdgpu_waves_per_eu(kW, 8)))
void k_batch_query(DsDev ds, BatchDev b)
{
    constexpr uint32_t kThreads = kW * 64;
    constexpr uint32_t kInSortMax = kThreads * kInItems; // result ids the block sorts itself
    static_assert(kInSortMax >= kResLds, "the in-block sort takes the whole LDS result buffer");
    static_assert((kW * kRadixDigits + kW) * sizeof(uint32_t) <= kCandBytes, "radix histograms in the candidate area");
    extern __shared__ __align__(16) unsigned char smem[];
    QueryCtl& ctl = *reinterpret_cast<QueryCtl*>(smem);
    uint32_t* ctl_words = reinterpret_cast<uint32_t*>(smem);
    uint32_t* cover = reinterpret_cast<uint32_t*>(smem + kCtlBytes);
    unsigned char* cand = smem + kCtlBytes + kCoverBytes;
    SharedRecsIdx<kRecLds>& S = *reinterpret_cast<SharedRecsIdx<kRecLds>*>(cand);
    WaveLds* wl = reinterpret_cast<WaveLds*>(cand);
    uint32_t* res = reinterpret_cast<uint32_t*>(cand + kCandBytes);
    uint32_t* front = reinterpret_cast<uint32_t*>(cand + kCandBytes + kResBytes);

    for (uint32_t i = threadIdx.x; i < kPhaseItems * 64; i += kThreads) cover[i] = 0;
    for (uint32_t qi = blockIdx.x; qi < b.nq; qi += gridDim.x) {
        // per query: the thread's addresses are formed afresh
        const uint32_t tid = opaque_u32(threadIdx.x), wave = tid >> 6, lane = opaque_u32(lane_id());
        const QDev Q = uniform_q(b.q[qi]);
        uint32_t* stat = b.qstat + (uint64_t)qi * kQueryStats;
        if (Q.trivial) { // k > |C|: every user qualifies; the host fills the result
            if (tid < kQueryStats) stat[tid] = 0;
            if (tid == 0) {
                b.qcount[qi] = 0;
                b.qland[qi] = 0;
                put_words(b, qi, 0, 0, Q);
            }
            continue;
        }
        if (tid < kCtlWords) ctl_words[tid] = 0;
        __syncthreads();
        // candidate records (binary64 -> fp32): into LDS sorted by sign
        // quadrant when they fit, each with its facility index in the spare
        // word, else into the batch's global record array
        const bool shared = Q.nc <= kRecLds;
        const double2 w = b.wbound[qi];
        if (shared) {
            constexpr int kPer = (kRecLds + kThreads - 1) / kThreads; // records per thread
            CandRec r[kPer];
            uint32_t qd[kPer];
#pragma unroll
            for (int h = 0; h < kPer; ++h) {
                const uint32_t j = tid + h * kThreads;
                r[h] = {0.f, 0.f, 0.f, 0.f};
                qd[h] = 4;
                if (j < Q.nc) {
                    const uint32_t f = b.cand[Q.off + j];
                    r[h] = make_record(Q.qx, Q.qy, b.fx[f], b.fy[f], w.x, w.y);
                    r[h].pad = __uint_as_float(f);
                    qd[h] = quadrant(r[h]);
                }
            }
            uint32_t rank[kPer];
#pragma unroll
            for (int h = 0; h < kPer; ++h) {
                rank[h] = 0;
                for (uint32_t g = 0; g < 4; ++g) { // wave-aggregated quadrant counts
                    const uint64_t mk = ballot(qd[h] == g);
                    if (qd[h] == g) rank[h] = rank_below(mk);
                    uint32_t pos = 0;
                    if (lane == 0 && mk) pos = atomicAdd(&ctl.qcnt[g], (uint32_t)__popcll(mk));
                    pos = uniform(pos);
                    if (qd[h] == g) rank[h] += pos;
                }
            }
            __syncthreads();
            uint32_t start[5];
            start[0] = 0;
            for (int g = 0; g < 4; ++g) start[g + 1] = start[g] + ctl.qcnt[g];
#pragma unroll
            for (int h = 0; h < kPer; ++h)
                if (qd[h] < 4) S.rec[start[qd[h]] + rank[h]] = r[h];
            if (tid < 5) ctl.quad[tid] = start[tid];
        } else {
            CandRec* grec = b.rec + Q.off;
            float2* gorig = b.orig + Q.off;
            for (uint32_t j = tid; j < Q.nc; j += kThreads) {
                const uint32_t f = b.cand[Q.off + j];
                const float x = b.fx[f], y = b.fy[f];
                grec[j] = make_record(Q.qx, Q.qy, x, y, w.x, w.y);
                gorig[j] = make_float2(x, y);
            }
        }
        __syncthreads();
        if (shared)
            descend<kW, true>(ds, b, Q, qi, ctl, front, cover, S.rec, nullptr, &wl[wave], res);
        else
            descend<kW, false>(ds, b, Q, qi, ctl, front, cover, b.rec + Q.off, b.orig + Q.off, &wl[wave], res);
        __syncthreads();
        const uint32_t nres = ctl.nres;
        if (tid == 0) {
            b.qcount[qi] = nres;
            stat[0] = ctl.fp64;
        }
        if (tid < kMaxLevels) stat[1 + tid] = ctl.items[tid];
        int32_t* seg = b.seg + (uint64_t)qi * b.seg_cap;
        uint32_t radix = 0; // the result goes through the radix sort
        uint32_t landed = kNotLanded;
        if (nres == 0) {
            landed = 0;
        } else if (nres <= b.seg_cap && nres <= kInSortMax) {
            // the keys: the LDS buffer's, then the spill in the segment
            uint32_t key[kInItems];
#pragma unroll
            for (int i = 0; i < kInItems; ++i) {
                const uint32_t j = tid + i * kThreads;
                key[i] = j < nres ? ((j < kResLds ? res[j] : (uint32_t)seg[j]) ^ 0x80000000u) : 0u;
            }
            __syncthreads(); // every key is in registers: the LDS from the cover counts on is free
            uint32_t* s_out = cover;
            if (bucket_sort_regs<kThreads, kInItems, kResBuckets>(key, nres, s_out, s_out + kInSortMax, ctl.red)) {
                if (tid == 0) ctl.land = reserve_landing(b, nres);
                __syncthreads();
                landed = ctl.land;
                // a full landing area: the sorted ids stay in the segment (the run is repeated)
                int32_t* dst = landed != kNotLanded ? b.hland + landed : seg;
                for (uint32_t i = tid; i < nres; i += kThreads) dst[i] = (int32_t)(s_out[i] ^ 0x80000000u);
            } else { // bunched ids: to the segment for the radix sort
#pragma unroll
                for (int i = 0; i < kInItems; ++i) {
                    const uint32_t j = tid + i * kThreads;
                    if (j < nres) seg[j] = (int32_t)(key[i] ^ 0x80000000u);
                }
                radix = 1;
            }
            // the cover counts start at zero for the next query
            __syncthreads();
            for (uint32_t i = tid; i < kPhaseItems * 64; i += kThreads) cover[i] = 0;
        } else if (nres <= b.seg_cap) { // the LDS part joins the spill in the segment
            for (uint32_t i = tid; i < kResLds; i += kThreads) seg[i] = (int32_t)res[i];
            radix = 1;
        }
        if (radix) {
            __syncthreads(); // the segment holds every id
            if (tid == 0) ctl.land = reserve_landing(b, nres);
            __syncthreads();
            landed = ctl.land;
            // a full landing area: the ids stay in the segment (the run is repeated)
            if (landed != kNotLanded) {
                // the radix histograms in the candidate area (free after the descent)
                auto* hist = reinterpret_cast<uint32_t(*)[kRadixDigits]>(cand);
                radix_land<kW>(b, seg, nres, landed, hist, reinterpret_cast<uint32_t*>(cand) + kW * kRadixDigits, tid,
                               wave, lane);
            }
        }
        if (tid == 0) {
            stat[kSortSlot] = radix;
            b.qland[qi] = landed;
            put_words(b, qi, nres, landed, Q);
        }
        __syncthreads(); // before the next query resets the state
    }
    // the last block to finish clears the reservation counters for the next
    // run (every reservation of this run is done by then)
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&b.land_ctr[1], 1ull) == (unsigned long long)gridDim.x - 1) {
            b.land_ctr[0] = 0;
            b.land_ctr[1] = 0;
        }
    }
}

int grid_for(uint64_t n, int per_block, int max_blocks)
{
    uint64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > (uint64_t)max_blocks) g = max_blocks;
    return (int)g;
}

} // namespace

int device_grid(int n_cu, int waves_per_cu) { return n_cu * waves_per_cu / kWaves; }

void launch_curve_keys(const float* ux, const float* uy, uint32_t n, float minx, float miny, float sx, float sy,
                       uint32_t* keys, uint32_t* idx, hipStream_t st)
{
    k_curve_keys<<<grid_for(n, 256, 4096), 256, 0, st>>>(ux, uy, n, minx, miny, sx, sy, keys, idx);
}

void launch_gather_users(const uint32_t* order, uint32_t n, uint32_t n_pad, const float* ux_in, const float* uy_in,
                         const int32_t* uid_in, float* ux, float* uy, int32_t* uid, hipStream_t st)
{
    k_gather_users<<<grid_for(n_pad, 256, 4096), 256, 0, st>>>(order, n, n_pad, ux_in, uy_in, uid_in, ux, uy, uid);
}

void launch_tile_mbr(const float* ux, const float* uy, uint32_t n_tiles, float4* mbr, hipStream_t st)
{
    k_tile_mbr<<<grid_for(n_tiles, kWaves, 4096), kBlock, 0, st>>>(ux, uy, n_tiles, mbr);
}

void launch_level_mbr(const float4* child, uint32_t n_child, float4* parent, uint32_t n_parent, hipStream_t st)
{
    k_level_mbr<<<grid_for(n_parent, kWaves, 4096), kBlock, 0, st>>>(child, n_child, parent, n_parent);
}

void launch_group_mbr(const float4* tiles, uint32_t n_tiles, uint32_t fan, float4* groups, hipStream_t st)
{
    k_group_mbr<<<grid_for((n_tiles + fan - 1) / fan, 256, 4096), 256, 0, st>>>(tiles, n_tiles, fan, groups);
}

void launch_plain(const DsDev& ds, QDev q, const CandRec* rec, const float2* orig, int32_t* out_ids,
                  uint32_t out_cap, uint32_t* counters, int grid, hipStream_t st)
{
    k_plain<<<grid_for(ds.n_tiles, kWaves, grid), kBlock, 0, st>>>(ds, q, rec, orig, out_ids, out_cap, counters);
}

void launch_build_records(const uint32_t* cand, uint32_t nc, uint32_t q, float qx, float qy, double wx, double wy,
                          const float* fx, const float* fy, CandRec* rec, float2* orig, uint32_t* ctr,
                          uint32_t n_ctr, hipStream_t st)
{
    k_build_records<<<grid_for(nc, 256, 64), 256, 0, st>>>(cand, nc, q, qx, qy, wx, wy, fx, fy, rec, orig, ctr,
                                                           n_ctr);
}

// Count-all grid: a block for every kWaves groups (capped; the kernel loops
// past the cap).
template <typename OutT>
void launch_count_typed(const DsDev& ds, QDev q, const CandRec* rec, const float2* orig, OutT* counts,
                        uint32_t* rechecks, int variant, hipStream_t st)
{
    const size_t lds = q.nc <= kCountLdsRecs ? q.nc * sizeof(CandRec) : 0;
    if (variant == kCountValu8) {
        const int g = grid_for((ds.n_tiles + 7) / 8, kWaves, 1 << 20);
        k_count_all<8, 1, OutT><<<g, kBlock, lds, st>>>(ds, q, rec, orig, counts, rechecks);
    } else if (variant == kCountValu6) {
        const int g = grid_for((ds.n_tiles + 5) / 6, kWaves, 1 << 20);
        k_count_all<6, 1, OutT><<<g, kBlock, lds, st>>>(ds, q, rec, orig, counts, rechecks);
    } else {
        const int g = grid_for((ds.n_tiles + 3) / 4, kWaves, 1 << 20);
        k_count_all<4, 1, OutT><<<g, kBlock, lds, st>>>(ds, q, rec, orig, counts, rechecks);
    }
}

template <typename OutT>
void launch_count_set_typed(const DsDev& ds, const CountSetQ* qs, uint32_t nq, uint32_t n_recs, const CandRec* rec,
                            const float2* orig, OutT* counts, uint32_t* rechecks, int variant, hipStream_t st)
{
    const size_t lds = (size_t)n_recs * sizeof(CandRec);
    if (variant == kCountValu8) {
        const int g = grid_for((ds.n_tiles + 7) / 8, kWaves, 1 << 20);
        k_count_set<8, OutT><<<g, kBlock, lds, st>>>(ds, qs, nq, n_recs, rec, orig, counts, rechecks);
    } else if (variant == kCountValu6) {
        const int g = grid_for((ds.n_tiles + 5) / 6, kWaves, 1 << 20);
        k_count_set<6, OutT><<<g, kBlock, lds, st>>>(ds, qs, nq, n_recs, rec, orig, counts, rechecks);
    } else {
        const int g = grid_for((ds.n_tiles + 3) / 4, kWaves, 1 << 20);
        k_count_set<4, OutT><<<g, kBlock, lds, st>>>(ds, qs, nq, n_recs, rec, orig, counts, rechecks);
    }
}

void launch_count_set(const DsDev& ds, const CountSetQ* qs, uint32_t nq, uint32_t n_recs, const CandRec* rec,
                      const float2* orig, void* counts, int bytes, uint32_t* rechecks, int variant, hipStream_t st)
{
    switch (bytes) {
    case 1: launch_count_set_typed(ds, qs, nq, n_recs, rec, orig, (uint8_t*)counts, rechecks, variant, st); break;
    case 2: launch_count_set_typed(ds, qs, nq, n_recs, rec, orig, (uint16_t*)counts, rechecks, variant, st); break;
    default: launch_count_set_typed(ds, qs, nq, n_recs, rec, orig, (uint32_t*)counts, rechecks, variant, st); break;
    }
}

const char* count_variant_name(int variant)
{
    static const char* const names[kCountVariants] = {"valu4", "valu8", "valu6"};
    return variant >= 0 && variant < kCountVariants ? names[variant] : "valu4";
}

int count_bytes(uint32_t nc) { return nc < 256u ? 1 : nc < 65536u ? 2 : 4; }

void launch_count_all(const DsDev& ds, QDev q, const CandRec* rec, const float2* orig, void* counts,
                      uint32_t* rechecks, int variant, hipStream_t st, bool wide)
{
    switch (wide ? 4 : count_bytes(q.nc)) {
    case 1: launch_count_typed(ds, q, rec, orig, (uint8_t*)counts, rechecks, variant, st); break;
    case 2: launch_count_typed(ds, q, rec, orig, (uint16_t*)counts, rechecks, variant, st); break;
    default: launch_count_typed(ds, q, rec, orig, (uint32_t*)counts, rechecks, variant, st); break;
    }
}

void launch_scatter_counts(const void* src_sorted, int bytes, const uint32_t* perm, uint32_t n,
                           uint32_t* dst_orig, hipStream_t st)
{
    const int g = grid_for(n, 256, 4096);
    switch (bytes) {
    case 1: k_scatter_u32<<<g, 256, 0, st>>>((const uint8_t*)src_sorted, perm, n, dst_orig); break;
    case 2: k_scatter_u32<<<g, 256, 0, st>>>((const uint16_t*)src_sorted, perm, n, dst_orig); break;
    default: k_scatter_u32<<<g, 256, 0, st>>>((const uint32_t*)src_sorted, perm, n, dst_orig); break;
    }
}

size_t batch_query_lds_bytes(uint32_t n_levels)
{
    return kCtlBytes + kCoverBytes + kCandBytes + kResBytes +
           ((size_t)(n_levels - 1) * kFront + 64) * sizeof(uint32_t);
}

void launch_batch_query(const DsDev& ds, const BatchDev& b, int shape, hipStream_t st)
{
    const uint32_t grid = b.nq < 65535 ? (b.nq ? b.nq : 1) : 65535;
    if (shape == kShape4Waves)
        k_batch_query<4><<<grid, 256, batch_query_lds_bytes(ds.n_levels), st>>>(ds, b);
    else
        k_batch_query<8><<<grid, 512, batch_query_lds_bytes(ds.n_levels), st>>>(ds, b);
}

} // namespace rk
