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
#include <hipcub/block/block_radix_sort.hpp>
#include <hipcub/hipcub.hpp>

#include "rk_common.hpp"
#include "rk_kernels.hpp"

namespace rk {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

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

// Wave-private LDS staging of one 64-candidate chunk.
struct WaveLds {
    CandRec rec[64];
    float2 orig[64];
};

// One query against one tile of 64 users.  Phase A puts candidates on lanes
// and classifies each against the tile's bounding rectangle: certainly
// covering the whole tile (counted once for every user), certainly missing
// it, or crossing it.  Phase B puts users on lanes and evaluates only the
// crossing candidates; a (user, candidate) test whose fp32 value lands in the
// band is settled on the spot by the oracle's binary64 expression, so the
// count is exact when the loop ends.
__device__ LeafOut leaf_tile(const DsDev& ds, float qx, float qy, uint32_t k, uint32_t nc,
                             const CandRec* __restrict__ rec, const float2* __restrict__ orig, uint32_t tile,
                             WaveLds* L)
{
    const uint32_t lane = lane_id();
    k = uniform(k);
    nc = uniform(nc);
    tile = uniform(tile);
    const uint32_t u = tile * kTile + lane;
    const bool valid = u < ds.nu;
    float ux = 0.f, uy = 0.f, wx = 0.f, wy = 0.f; // loaded on the first crossing candidate
    double dq = 0.0;
    bool loaded = false; // wave-uniform
    const float4 m = ds.mbr[0][tile];
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
            o = orig[j];
        }
        const float tlo = fmaf(r.gx >= 0.f ? lx : hx, r.gx, fmaf(r.gy >= 0.f ? ly : hy, r.gy, -r.sh));
        const float thi = fmaf(r.gx >= 0.f ? hx : lx, r.gx, fmaf(r.gy >= 0.f ? hy : ly, r.gy, -r.sh));
        const bool full = have && tlo >= kCertainHi;
        const bool cross = have && !full && thi > kCertainLo;
        base += (uint32_t)__popcll(__ballot(full));
        if (base >= k) {
            dead = true;
            break;
        }
        uint64_t cm = __ballot(cross);
        if (cm) {
            if (!loaded) { // culled tiles never touch their users (arrays padded to whole tiles)
                ux = ds.ux[u];
                uy = ds.uy[u];
                wx = ux - qx;
                wy = uy - qy;
                dq = d2_exact(ux, uy, qx, qy);
                loaded = true;
            }
            L->rec[lane] = r;
            L->orig[lane] = o;
            __builtin_amdgcn_wave_barrier();
            while (cm) {
                const int b = __ffsll((unsigned long long)cm) - 1;
                cm &= cm - 1;
                const CandRec c = L->rec[b];
                const float t = fmaf(wx, c.gx, fmaf(wy, c.gy, -c.sh));
                cnt += (t >= kCertainHi) ? 1u : 0u;
                const bool inband = (t > kCertainLo) && (t < kCertainHi);
                const uint64_t bm = __ballot(inband);
                if (bm) { // rare: settle the band in binary64
                    const float2 f = L->orig[b];
                    if (inband && d2_exact(ux, uy, f.x, f.y) < dq) ++cnt;
                    fp64 += (uint32_t)__popcll(bm & __ballot(valid));
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (__ballot(valid && base + cnt < k) == 0) {
            dead = true;
            break;
        }
    }
    LeafOut out;
    out.keep = __ballot(!dead && valid && base + cnt < k);
    out.fp64 = fp64;
    return out;
}

// ---------------------------------------------------------------- plain form

__global__ __launch_bounds__(kBlock) void k_plain(DsDev ds, QDev q, const CandRec* __restrict__ rec,
                                                  const float2* __restrict__ orig, int32_t* __restrict__ out,
                                                  uint32_t* __restrict__ ctr)
{
    __shared__ WaveLds s_w[kWaves];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    const uint32_t stride = gridDim.x * kWaves;
    uint32_t rech = 0;
    for (uint32_t tile = blockIdx.x * kWaves + wave; tile < ds.n_tiles; tile += stride) {
        const LeafOut r = leaf_tile(ds, q.qx, q.qy, q.k, q.nc, rec, orig, tile, &s_w[wave]);
        rech += r.fp64;
        if (r.keep) {
            uint32_t pos = 0;
            if (lane == 0) pos = atomicAdd(&ctr[0], (uint32_t)__popcll(r.keep));
            pos = uniform(pos);
            if ((r.keep >> lane) & 1ull) out[pos + rank_below(r.keep)] = ds.uid[tile * kTile + lane];
        }
    }
    if (lane == 0 && rech) atomicAdd(&ctr[1 + (blockIdx.x * kWaves + wave) % kBuckets], rech);
}

// ------------------------------------------------------------ count-all form

// Every (user, candidate) pair is evaluated: no culling, no early exit.  A
// wave owns 4 consecutive tiles (256 users, 4 per lane) so that each
// broadcast LDS read of a record feeds 4 users and the loop is VALU-bound.
// Per 64-candidate chunk the candidates are split against the group's
// rectangle: "clean" ones, whose computed t is >= 19/16 (or <= -3/16) at the
// rectangle's extreme corner, give every user t >= 1 (or <= 0) because the
// fp32 error is < 1/16 per evaluation, so the saturated value is exactly 1
// or 0 and a float sum counts them exactly (2 FMAs with clamp + 1 add per
// test); the others take the banded path with the binary64 re-check.
constexpr int kCountUsers = 4;

__global__ __launch_bounds__(kBlock) void k_count_all(DsDev ds, QDev q, const CandRec* __restrict__ rec,
                                                      const float2* __restrict__ orig,
                                                      uint32_t* __restrict__ counts, uint32_t* __restrict__ rech_out)
{
    __shared__ CandRec s_rec[kWaves][128];
    __shared__ float2 s_orig[kWaves][64];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    const uint32_t stride = gridDim.x * kWaves;
    const uint32_t n_groups = (ds.n_tiles + kCountUsers - 1) / kCountUsers;
    uint32_t rech = 0;
    CandRec* wlds = s_rec[wave];
    float2* worig = s_orig[wave];
    for (uint32_t g = blockIdx.x * kWaves + wave; g < n_groups; g += stride) {
        const uint32_t t0 = g * kCountUsers;
        const uint32_t nt = min((uint32_t)kCountUsers, ds.n_tiles - t0);
        float4 m = ds.mbr[0][t0];
        float px[kCountUsers], py[kCountUsers], wx[kCountUsers], wy[kCountUsers], acc[kCountUsers];
        double dq[kCountUsers];
        uint32_t cnt[kCountUsers];
#pragma unroll
        for (int j = 0; j < kCountUsers; ++j) {
            px[j] = py[j] = wx[j] = wy[j] = acc[j] = 0.f;
            dq[j] = 0.0;
            cnt[j] = 0;
            if ((uint32_t)j < nt) {
                const uint32_t u = (t0 + j) * kTile + lane;
                px[j] = ds.ux[u];
                py[j] = ds.uy[u];
                wx[j] = px[j] - q.qx;
                wy[j] = py[j] - q.qy;
                dq[j] = d2_exact(px[j], py[j], q.qx, q.qy);
                if (j) {
                    const float4 b = ds.mbr[0][t0 + j];
                    m.x = fminf(m.x, b.x);
                    m.y = fminf(m.y, b.y);
                    m.z = fmaxf(m.z, b.z);
                    m.w = fmaxf(m.w, b.w);
                }
            }
        }
        const float lx = m.x - q.qx, ly = m.y - q.qy, hx = m.z - q.qx, hy = m.w - q.qy;
        for (uint32_t c0 = 0; c0 < q.nc; c0 += 64) {
            const uint32_t jc = c0 + lane;
            const bool have = jc < q.nc;
            CandRec r = {0.f, 0.f, 0.f, 0.f};
            if (have) r = rec[jc];
            const float tlo = fmaf(r.gx >= 0.f ? lx : hx, r.gx, fmaf(r.gy >= 0.f ? ly : hy, r.gy, -r.sh));
            const float thi = fmaf(r.gx >= 0.f ? hx : lx, r.gx, fmaf(r.gy >= 0.f ? hy : ly, r.gy, -r.sh));
            const bool clean = have && (tlo >= 1.1875f || thi <= -0.1875f);
            const bool slow = have && !clean;
            const uint64_t mc = __ballot(clean), ms = __ballot(slow);
            if (clean) wlds[rank_below(mc)] = r;
            if (slow) {
                wlds[64 + rank_below(ms)] = r;
                worig[rank_below(ms)] = orig[jc];
            }
            __builtin_amdgcn_wave_barrier();
            const uint32_t n_clean = (uint32_t)__popcll(mc), n_slow = (uint32_t)__popcll(ms);
#pragma unroll 4
            for (uint32_t i = 0; i < n_clean; ++i) {
                const CandRec c = wlds[i];
#pragma unroll
                for (int j = 0; j < kCountUsers; ++j) // med3(t,0,1) folds into the FMA's clamp bit
                    acc[j] += __builtin_amdgcn_fmed3f(fmaf(wx[j], c.gx, fmaf(wy[j], c.gy, -c.sh)), 0.0f, 1.0f);
            }
            for (uint32_t i = 0; i < n_slow; ++i) {
                const CandRec c = wlds[64 + i];
#pragma unroll
                for (int j = 0; j < kCountUsers; ++j) {
                    const float t = fmaf(wx[j], c.gx, fmaf(wy[j], c.gy, -c.sh));
                    cnt[j] += (t >= kCertainHi) ? 1u : 0u;
                    const bool inband = (t > kCertainLo) && (t < kCertainHi);
                    const uint64_t bm = __ballot(inband);
                    if (bm) { // rare: settle the band in binary64
                        const float2 f = worig[i];
                        if (inband && d2_exact(px[j], py[j], f.x, f.y) < dq[j]) ++cnt[j];
                        rech += (uint32_t)__popcll(bm);
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int j = 0; j < kCountUsers; ++j) {
            if ((uint32_t)j >= nt) break;
            counts[(t0 + j) * kTile + lane] = (uint32_t)acc[j] + cnt[j];
        }
    }
    if (lane == 0 && rech) atomicAdd(&rech_out[(blockIdx.x * kWaves + wave) % kBuckets], rech);
}

__global__ void k_scatter_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ perm, uint32_t n,
                              uint32_t* __restrict__ dst)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[perm[i]] = src[i];
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

// -------------------------------------------------------------- batched form

// First kernel of a batch step: zero the counters, clear each query's
// result count and build its candidate records (one block per query).
__global__ void k_batch_prep(BatchDev b, const uint32_t* __restrict__ cand, const float* __restrict__ fx,
                             const float* __restrict__ fy, const double2* __restrict__ wb, uint64_t total)
{
    if (blockIdx.x == 0) {
        uint32_t* c = reinterpret_cast<uint32_t*>(b.ctr);
        for (uint32_t i = threadIdx.x; i < sizeof(BatchCounters) / 4; i += blockDim.x) c[i] = 0;
    }
    for (uint32_t qi = blockIdx.x; qi < b.nq; qi += gridDim.x) {
        const QDev Q = b.q[qi];
        if (threadIdx.x == 0) b.qcount[qi] = 0;
        const double2 w = wb[qi];
        for (uint32_t j = threadIdx.x; j < Q.nc; j += blockDim.x) {
            const uint64_t g = Q.off + j;
            if (g >= total) break;
            const uint32_t f = cand[g];
            const float x = fx[f], y = fy[f];
            const_cast<CandRec*>(b.rec)[g] = make_record(Q.qx, Q.qy, x, y, w.x, w.y);
            const_cast<float2*>(b.orig)[g] = make_float2(x, y);
        }
    }
}

// Items at parent level p are (query, node); the node's <= 64 children at
// level p-1 go on lanes.  A child is culled once k candidates certainly cover
// its whole rectangle (the internal-node test of rknn-inf/inf.h:219-237,
// with certified corner evaluation instead of a 4-corner FP64 test).
// Survivors are staged in a wave-private LDS buffer and appended to level
// p-1 with one atomic per flush instead of one per item.
constexpr uint32_t kStage = 192;

// Per-wave view of a bucketed work list: lane b holds bucket b's exclusive
// and inclusive prefix of the append counts (clamped to the bucket size).
struct ListView {
    const uint64_t* base;
    uint32_t bcap;
    uint32_t excl, incl, total;
};

__device__ ListView list_view(const uint64_t* base, const uint32_t* counts, uint32_t bcap)
{
    const uint32_t lane = lane_id();
    const uint32_t c = min(counts[lane], bcap);
    uint32_t incl = c;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)incl, o);
        if ((int)lane >= o) incl += v;
    }
    ListView v{base, bcap, incl - c, incl, 0};
    v.total = uniform((uint32_t)__builtin_amdgcn_readlane((int)incl, 63));
    return v;
}

// Item i of the list (all lanes call it with the same i < total).
__device__ uint64_t list_item(const ListView& v, uint32_t i)
{
    const uint32_t bucket = uniform((uint32_t)__popcll(__ballot(v.incl <= i)));
    const uint32_t start = (uint32_t)__builtin_amdgcn_readlane((int)v.excl, (int)bucket);
    return uniform64(v.base[(uint64_t)bucket * v.bcap + (i - start)]);
}

__device__ void flush_items(uint64_t* stage, uint32_t& n, uint32_t* counters, uint64_t* next, uint32_t bcap,
                            uint32_t* overflow)
{
    if (n == 0) return;
    const uint32_t lane = lane_id();
    const uint32_t bucket = (blockIdx.x * kWaves + (threadIdx.x >> 6)) % kBuckets;
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(&counters[bucket], n);
    pos = uniform(pos);
    __builtin_amdgcn_wave_barrier();
    uint64_t* dst = next + (uint64_t)bucket * bcap;
    for (uint32_t i = lane; i < n; i += 64) {
        if (pos + i < bcap)
            dst[pos + i] = stage[i];
        else
            *overflow = 1;
    }
    __builtin_amdgcn_wave_barrier();
    n = 0;
}

// Lanes = the <= 64 children of `node` at a level with MBRs `cm`: the mask of
// children that fewer than k candidates certainly cover.  Candidate records
// are staged 64 at a time in wlds; `staged` says chunk 0 is already there.
__device__ uint64_t surviving_children(const float4* __restrict__ cm, uint32_t n_child, uint32_t node,
                                       const QDev& Q, const CandRec* __restrict__ rec, CandRec* wlds, bool staged)
{
    const uint32_t lane = lane_id();
    const uint32_t c = node * kFanout + lane;
    const bool valid = c < n_child;
    const float4 m = cm[valid ? c : 0];
    const float lx = m.x - Q.qx, ly = m.y - Q.qy, hx = m.z - Q.qx, hy = m.w - Q.qy;
    uint32_t full = 0;
    for (uint32_t c0 = 0; c0 < Q.nc; c0 += 64) {
        const uint32_t n = min(64u, Q.nc - c0);
        if (c0 > 0 || !staged) {
            __builtin_amdgcn_wave_barrier();
            if (lane < n) wlds[lane] = rec[c0 + lane];
            __builtin_amdgcn_wave_barrier();
        }
        bool live = true;
        for (uint32_t i0 = 0; i0 < n && live; i0 += 16) {
            if (i0 + 16 <= n) { // full block: unrolled so the LDS reads pipeline
#pragma unroll
                for (uint32_t i = 0; i < 16; ++i) {
                    const CandRec r = wlds[i0 + i];
                    const float t = fmaf(r.gx >= 0.f ? lx : hx, r.gx, fmaf(r.gy >= 0.f ? ly : hy, r.gy, -r.sh));
                    full += (t >= kCertainHi) ? 1u : 0u;
                }
            } else {
                for (uint32_t i = i0; i < n; ++i) {
                    const CandRec r = wlds[i];
                    const float t = fmaf(r.gx >= 0.f ? lx : hx, r.gx, fmaf(r.gy >= 0.f ? ly : hy, r.gy, -r.sh));
                    full += (t >= kCertainHi) ? 1u : 0u;
                }
            }
            live = __ballot(valid && full < Q.k) != 0;
        }
        if (!live) return 0;
    }
    return __ballot(valid && full < Q.k);
}

__device__ __forceinline__ void stage_chunk0(const QDev& Q, const CandRec* __restrict__ rec, CandRec* wlds)
{
    const uint32_t lane = lane_id();
    __builtin_amdgcn_wave_barrier();
    if (lane < min(64u, Q.nc)) wlds[lane] = rec[lane];
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void push_children(uint64_t surv, uint32_t qi, uint32_t node, uint64_t* stage,
                                              uint32_t& staged, uint32_t* counters, uint64_t* next, uint32_t bcap,
                                              uint32_t* overflow)
{
    const uint32_t ns = (uint32_t)__popcll(surv);
    if (!ns) return;
    if (staged + ns > kStage) flush_items(stage, staged, counters, next, bcap, overflow);
    const uint32_t lane = lane_id();
    if ((surv >> lane) & 1ull) stage[staged + rank_below(surv)] = ((uint64_t)qi << 32) | (node * kFanout + lane);
    staged += ns;
    __builtin_amdgcn_wave_barrier();
}

// Root of every query: one wave per query tests the <= 64 top tree nodes
// and lists the survivors for the level kernels.
__global__ __launch_bounds__(kBlock) void k_batch_root(DsDev ds, BatchDev b)
{
    __shared__ CandRec s_rec[kWaves][64];
    __shared__ uint64_t s_out[kWaves][kStage];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t stride = gridDim.x * kWaves;
    const uint32_t top = ds.n_levels - 1; // level of the <= 64 top nodes
    uint64_t* __restrict__ next = b.items[top];
    const uint32_t bcap = b.item_cap[top];
    CandRec* wlds = s_rec[wave];
    uint64_t* stage = s_out[wave];
    uint32_t staged = 0;
    for (uint32_t qi = blockIdx.x * kWaves + wave; qi < b.nq; qi += stride) {
        const QDev Q = uniform_q(b.q[qi]);
        if (Q.trivial) continue;
        const CandRec* __restrict__ rec = b.rec + Q.off;
        stage_chunk0(Q, rec, wlds);
        const uint64_t surv = surviving_children(ds.mbr[top], ds.nlev[top], 0, Q, rec, wlds, true);
        push_children(surv, qi, 0, stage, staged, b.ctr->items[top], next, bcap, &b.ctr->overflow);
    }
    flush_items(stage, staged, b.ctr->items[top], next, bcap, &b.ctr->overflow);
}

// Items at parent level p: (query, node at level p); survivors among the
// node's children are pushed to level p-1.
__global__ __launch_bounds__(kBlock) void k_batch_internal(DsDev ds, BatchDev b, uint32_t p)
{
    __shared__ CandRec s_rec[kWaves][64];
    __shared__ uint64_t s_out[kWaves][kStage];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t stride = gridDim.x * kWaves;
    const ListView lv = list_view(b.items[p], b.ctr->items[p], b.item_cap[p]);
    const uint32_t n_items = lv.total;
    const uint32_t cl = p - 1;
    uint64_t* __restrict__ next = b.items[cl];
    const uint32_t next_cap = b.item_cap[cl];
    CandRec* wlds = s_rec[wave];
    uint64_t* stage = s_out[wave];
    uint32_t staged = 0; // wave-uniform
    uint32_t it = blockIdx.x * kWaves + wave;
    uint64_t item = it < n_items ? list_item(lv, it) : 0;
    for (; it < n_items; it += stride) {
        const uint64_t nxt = it + stride < n_items ? list_item(lv, it + stride) : 0; // prefetch
        const uint32_t qi = uniform((uint32_t)(item >> 32)), node = uniform((uint32_t)item);
        const QDev Q = uniform_q(b.q[qi]);
        const CandRec* __restrict__ rec = b.rec + Q.off;
        const uint64_t surv = surviving_children(ds.mbr[cl], ds.nlev[cl], node, Q, rec, wlds, false);
        push_children(surv, qi, node, stage, staged, b.ctr->items[cl], next, next_cap, &b.ctr->overflow);
        item = nxt;
    }
    flush_items(stage, staged, b.ctr->items[cl], next, next_cap, &b.ctr->overflow);
}

// Leaf items (query, tile): the plain form's tile routine; survivors go to
// the query's own result segment (one atomic per item with survivors).
__global__ __launch_bounds__(kBlock) void k_batch_leaf(DsDev ds, BatchDev b)
{
    __shared__ WaveLds s_w[kWaves];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    const uint32_t stride = gridDim.x * kWaves;
    const ListView lv = list_view(b.items[0], b.ctr->items[0], b.item_cap[0]);
    const uint32_t n_items = lv.total;
    uint32_t rech = 0;
    uint32_t it = blockIdx.x * kWaves + wave;
    uint64_t item = it < n_items ? list_item(lv, it) : 0;
    for (; it < n_items; it += stride) {
        const uint64_t nxt = it + stride < n_items ? list_item(lv, it + stride) : 0; // prefetch
        const uint32_t qi = uniform((uint32_t)(item >> 32)), tile = uniform((uint32_t)item);
        const QDev Q = uniform_q(b.q[qi]);
        const LeafOut r = leaf_tile(ds, Q.qx, Q.qy, Q.k, Q.nc, b.rec + Q.off, b.orig + Q.off, tile, &s_w[wave]);
        rech += r.fp64;
        if (r.keep) {
            uint32_t rank0 = 0;
            if (lane == 0) rank0 = atomicAdd(&b.qcount[qi], (uint32_t)__popcll(r.keep));
            rank0 = uniform(rank0);
            if ((r.keep >> lane) & 1ull) {
                const uint32_t rk = rank0 + rank_below(r.keep);
                if (rk < b.seg_cap)
                    b.seg[(uint64_t)qi * b.seg_cap + rk] = ds.uid[tile * kTile + lane];
                else
                    b.ctr->overflow = 1;
            }
        }
        item = nxt;
    }
    if (lane == 0 && rech) atomicAdd(&b.ctr->rechecks[(blockIdx.x * kWaves + wave) % kBuckets], rech);
}

// Exclusive scan of per-query result counts (one block) for the compacted
// layout; also the batch total.
__global__ __launch_bounds__(1024) void k_batch_scan(BatchDev b)
{
    __shared__ uint64_t s_part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (b.nq + 1023) / 1024;
    const uint32_t lo = min(b.nq, t * per), hi = min(b.nq, lo + per);
    uint64_t sum = 0;
    for (uint32_t i = lo; i < hi; ++i) sum += min(b.qcount[i], b.seg_cap);
    s_part[t] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint64_t v = t >= o ? s_part[t - o] : 0;
        __syncthreads();
        s_part[t] += v;
        __syncthreads();
    }
    uint64_t run = s_part[t] - sum;
    for (uint32_t i = lo; i < hi; ++i) {
        b.qoff[i] = run;
        run += min(b.qcount[i], b.seg_cap);
    }
    if (t == 1023) {
        b.qoff[b.nq] = s_part[1023];
        b.ctr->results = (uint32_t)min<uint64_t>(s_part[1023], 0xffffffffull);
    }
}

// One block per query: radix sort of its result segment (rocPRIM block
// radix sort via hipCUB, keys held in registers).  Segments of up to 2048
// ids use 8 keys per thread, up to 4096 use 16; larger ones are flagged and
// sorted by the host when fetched.
using SortSmall = hipcub::BlockRadixSort<int32_t, 256, 8>;
using SortLarge = hipcub::BlockRadixSort<int32_t, 256, 16>;

template <class Sorter, int kItems>
__device__ void sort_segment(int32_t* seg, uint32_t n, typename Sorter::TempStorage& tmp)
{
    int32_t keys[kItems];
    const uint32_t base = threadIdx.x * kItems;
#pragma unroll
    for (int i = 0; i < kItems; ++i) keys[i] = base + i < n ? seg[base + i] : INT32_MAX;
    Sorter(tmp).Sort(keys);
#pragma unroll
    for (int i = 0; i < kItems; ++i)
        if (base + i < n) seg[base + i] = keys[i];
}

__global__ __launch_bounds__(256) void k_batch_sort(BatchDev b)
{
    __shared__ union {
        typename SortSmall::TempStorage small;
        typename SortLarge::TempStorage large;
    } tmp;
    for (uint32_t qi = blockIdx.x; qi < b.nq; qi += gridDim.x) {
        const uint32_t n = min(b.qcount[qi], b.seg_cap);
        int32_t* seg = b.seg + (uint64_t)qi * b.seg_cap;
        if (n <= 1) continue;
        if (n <= 256 * 8) {
            sort_segment<SortSmall, 8>(seg, n, tmp.small);
        } else if (n <= 256 * 16) {
            sort_segment<SortLarge, 16>(seg, n, tmp.large);
        } else if (threadIdx.x == 0) {
            atomicAdd(&b.ctr->sort_overflow, 1u);
        }
        __syncthreads();
    }
}

// Fetch-time: copy each query's segment to its place in the compacted array.
__global__ void k_batch_compact(BatchDev b)
{
    for (uint32_t qi = blockIdx.x; qi < b.nq; qi += gridDim.x) {
        const uint32_t n = min(b.qcount[qi], b.seg_cap);
        const int32_t* src = b.seg + (uint64_t)qi * b.seg_cap;
        int32_t* dst = b.out + b.qoff[qi];
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
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

void launch_plain(const DsDev& ds, QDev q, const CandRec* rec, const float2* orig, int32_t* out_ids,
                  uint32_t* counters, int grid, hipStream_t st)
{
    k_plain<<<grid_for(ds.n_tiles, kWaves, grid), kBlock, 0, st>>>(ds, q, rec, orig, out_ids, counters);
}

void launch_count_all(const DsDev& ds, QDev q, const CandRec* rec, const float2* orig, uint32_t* counts,
                      uint32_t* rechecks, int grid, hipStream_t st)
{
    const uint32_t groups = (ds.n_tiles + kCountUsers - 1) / kCountUsers;
    k_count_all<<<grid_for(groups, kWaves, grid), kBlock, 0, st>>>(ds, q, rec, orig, counts, rechecks);
}

void launch_scatter_u32(const uint32_t* src_sorted, const uint32_t* perm, uint32_t n, uint32_t* dst_orig,
                        hipStream_t st)
{
    k_scatter_u32<<<grid_for(n, 256, 4096), 256, 0, st>>>(src_sorted, perm, n, dst_orig);
}

void launch_batch_prep(const BatchDev& b, const uint32_t* cand, const float* fx, const float* fy,
                       const double2* wbound, uint64_t total, hipStream_t st)
{
    k_batch_prep<<<b.nq < 4096 ? (b.nq ? b.nq : 1) : 4096, 256, 0, st>>>(b, cand, fx, fy, wbound, total);
}

void launch_batch_root(const DsDev& ds, const BatchDev& b, int grid, hipStream_t st)
{
    k_batch_root<<<grid, kBlock, 0, st>>>(ds, b);
}

void launch_batch_internal(const DsDev& ds, const BatchDev& b, uint32_t parent_level, int grid, hipStream_t st)
{
    k_batch_internal<<<grid, kBlock, 0, st>>>(ds, b, parent_level);
}

void launch_batch_leaf(const DsDev& ds, const BatchDev& b, int grid, hipStream_t st)
{
    k_batch_leaf<<<grid, kBlock, 0, st>>>(ds, b);
}

void launch_batch_sort(const BatchDev& b, hipStream_t st)
{
    k_batch_sort<<<b.nq < 65535 ? (b.nq ? b.nq : 1) : 65535, 256, 0, st>>>(b);
}

void launch_batch_scan(const BatchDev& b, hipStream_t st)
{
    k_batch_scan<<<1, 1024, 0, st>>>(b);
}

void launch_batch_compact(const BatchDev& b, hipStream_t st)
{
    k_batch_compact<<<b.nq < 65535 ? (b.nq ? b.nq : 1) : 65535, 256, 0, st>>>(b);
}

} // namespace rk
