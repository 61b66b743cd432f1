// rk_kernels.hip -- gfx950 kernels of the RkNN dominance-count engine.
//
// Wave64 throughout: a leaf tile is 64 Morton-adjacent users (one per lane),
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
#include <hipcub/hipcub.hpp>

#include "rk_common.hpp"
#include "rk_kernels.hpp"

namespace rk {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kSortCap = 8192; // result ids sorted in LDS per query

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t rank_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_bcast(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src); }

// Exact count of candidates strictly closer than q under the oracle's
// binary64 expression (rknn-naive/rknn.cpp:45-47 distances, strict '<' of
// rknn-inf/inf.h:77-93), stopping at cap.
__device__ uint32_t exact_count(float ux, float uy, float qx, float qy, const float2* __restrict__ orig,
                                uint32_t nc, uint32_t cap)
{
    const double dq = d2_exact(ux, uy, qx, qy);
    uint32_t c = 0;
    for (uint32_t j = 0; j < nc; ++j) {
        const float2 f = orig[j];
        if (d2_exact(ux, uy, f.x, f.y) < dq) {
            if (++c >= cap) break;
        }
    }
    return c;
}

struct LeafOut {
    uint64_t keep;     // lanes (users) of the tile in R(q,k)
    uint32_t rechecks; // lanes resolved in binary64
};

// One query against one tile of 64 users.  Phase A puts candidates on lanes
// and classifies each against the tile's bounding rectangle: certainly
// covering the whole tile (counted once for every user), certainly missing
// it, or crossing it.  Phase B puts users on lanes and evaluates only the
// crossing candidates.  Users whose certain count stays below k but who hit
// the band of some candidate are re-counted in binary64.
__device__ LeafOut leaf_tile(const DsDev& ds, float qx, float qy, uint32_t k, uint32_t nc,
                             const CandRec* __restrict__ rec, const float2* __restrict__ orig, uint32_t tile,
                             CandRec* wlds)
{
    const uint32_t lane = lane_id();
    const uint32_t u = tile * kTile + lane;
    const bool valid = u < ds.nu;
    float wx = 0.f, wy = 0.f; // user offsets, loaded on the first crossing candidate
    bool loaded = false;      // wave-uniform
    const float4 m = ds.mbr[0][tile];
    const float lx = m.x - qx, ly = m.y - qy, hx = m.z - qx, hy = m.w - qy;

    uint32_t base = 0; // wave-uniform: candidates covering the whole tile
    uint32_t cnt = 0;  // per user: crossing candidates certainly closer
    bool band = false;
    bool dead = false; // every user already has >= k certain
    for (uint32_t c0 = 0; c0 < nc; c0 += 64) {
        const uint32_t j = c0 + lane;
        const bool have = j < nc;
        CandRec r = {0.f, 0.f, 0.f, 0.f};
        if (have) r = rec[j];
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
                wx = ds.ux[u] - qx;
                wy = ds.uy[u] - qy;
                loaded = true;
            }
            wlds[lane] = r;
            __builtin_amdgcn_wave_barrier();
            while (cm) {
                const int b = __ffsll((unsigned long long)cm) - 1;
                cm &= cm - 1;
                const CandRec c = wlds[b];
                const float t = fmaf(wx, c.gx, fmaf(wy, c.gy, -c.sh));
                cnt += (t >= kCertainHi) ? 1u : 0u;
                band |= (t > kCertainLo) && (t < kCertainHi);
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (__ballot(valid && base + cnt < k) == 0) {
            dead = true;
            break;
        }
    }
    LeafOut o;
    const bool alive = !dead && valid && (base + cnt < k);
    bool keep = alive;
    const bool need = alive && band;
    if (need) keep = exact_count(ds.ux[u], ds.uy[u], qx, qy, orig, nc, k) < k;
    o.keep = __ballot(keep);
    o.rechecks = (uint32_t)__popcll(__ballot(need));
    return o;
}

// ---------------------------------------------------------------- plain form

__global__ __launch_bounds__(kBlock) void k_plain(DsDev ds, QDev q, const CandRec* __restrict__ rec,
                                                  const float2* __restrict__ orig, int32_t* __restrict__ out,
                                                  uint32_t* __restrict__ ctr)
{
    __shared__ CandRec s_rec[kWaves][64];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    const uint32_t stride = gridDim.x * kWaves;
    uint32_t rech = 0;
    for (uint32_t tile = blockIdx.x * kWaves + wave; tile < ds.n_tiles; tile += stride) {
        const LeafOut r = leaf_tile(ds, q.qx, q.qy, q.k, q.nc, rec, orig, tile, s_rec[wave]);
        rech += r.rechecks;
        if (r.keep) {
            uint32_t pos = 0;
            if (lane == 0) pos = atomicAdd(&ctr[0], (uint32_t)__popcll(r.keep));
            pos = wave_bcast(pos, 0);
            if ((r.keep >> lane) & 1ull) out[pos + rank_below(r.keep)] = ds.uid[tile * kTile + lane];
        }
    }
    if (lane == 0 && rech) atomicAdd(&ctr[1], rech);
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
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    const uint32_t stride = gridDim.x * kWaves;
    const uint32_t n_groups = (ds.n_tiles + kCountUsers - 1) / kCountUsers;
    uint32_t rech = 0;
    CandRec* wlds = s_rec[wave];
    for (uint32_t g = blockIdx.x * kWaves + wave; g < n_groups; g += stride) {
        const uint32_t t0 = g * kCountUsers;
        const uint32_t nt = min((uint32_t)kCountUsers, ds.n_tiles - t0);
        float4 m = ds.mbr[0][t0];
        float wx[kCountUsers], wy[kCountUsers], acc[kCountUsers];
        uint32_t cnt[kCountUsers];
        bool band[kCountUsers];
#pragma unroll
        for (int j = 0; j < kCountUsers; ++j) {
            wx[j] = wy[j] = acc[j] = 0.f;
            cnt[j] = 0;
            band[j] = false;
            if ((uint32_t)j < nt) {
                const uint32_t u = (t0 + j) * kTile + lane;
                wx[j] = ds.ux[u] - q.qx;
                wy[j] = ds.uy[u] - q.qy;
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
            if (slow) wlds[64 + rank_below(ms)] = r;
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
                    band[j] |= (t > kCertainLo) && (t < kCertainHi);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int j = 0; j < kCountUsers; ++j) {
            if ((uint32_t)j >= nt) break;
            const uint32_t u = (t0 + j) * kTile + lane;
            uint32_t c = (uint32_t)acc[j] + cnt[j];
            if (band[j]) c = exact_count(ds.ux[u], ds.uy[u], q.qx, q.qy, orig, q.nc, 0xffffffffu);
            rech += (uint32_t)__popcll(__ballot(band[j] && u < ds.nu));
            counts[u] = c;
        }
    }
    if (lane == 0 && rech) atomicAdd(rech_out, rech);
}

__global__ void k_scatter_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ perm, uint32_t n,
                              uint32_t* __restrict__ dst)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[perm[i]] = src[i];
}

// ------------------------------------------------------ dataset construction

__device__ __forceinline__ uint32_t spread16(uint32_t v)
{
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__global__ void k_morton(const float* __restrict__ ux, const float* __restrict__ uy, uint32_t n, float minx,
                         float miny, float sx, float sy, uint32_t* __restrict__ keys, uint32_t* __restrict__ idx)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float fx = (ux[i] - minx) * sx, fy = (uy[i] - miny) * sy;
        uint32_t ix = (uint32_t)fminf(fmaxf(fx, 0.f), 65535.f);
        uint32_t iy = (uint32_t)fminf(fmaxf(fy, 0.f), 65535.f);
        keys[i] = spread16(ix) | (spread16(iy) << 1);
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

__global__ void k_prep_records(const QDev* __restrict__ qs, uint32_t nq, const uint32_t* __restrict__ cand,
                               const float* __restrict__ fx, const float* __restrict__ fy,
                               const double2* __restrict__ wb, CandRec* __restrict__ rec, float2* __restrict__ orig,
                               uint64_t total)
{
    // one block per query, threads over its candidates
    for (uint32_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
        const QDev Q = qs[qi];
        const double2 w = wb[qi];
        for (uint32_t j = threadIdx.x; j < Q.nc; j += blockDim.x) {
            const uint64_t g = Q.off + j;
            if (g >= total) break;
            const uint32_t f = cand[g];
            const float x = fx[f], y = fy[f];
            rec[g] = make_record(Q.qx, Q.qy, x, y, w.x, w.y);
            orig[g] = make_float2(x, y);
        }
    }
}

__global__ void k_batch_init(BatchDev b, uint32_t root_level)
{
    __shared__ uint32_t s_n;
    if (threadIdx.x == 0) {
        for (int l = 0; l <= kMaxLevels; ++l) b.ctr->items[l] = 0;
        b.ctr->overflow = 0;
        b.ctr->pool = 0;
        b.ctr->sort_overflow = 0;
        b.ctr->rechecks = 0;
        s_n = 0;
    }
    __syncthreads();
    for (uint32_t qi = threadIdx.x; qi < b.nq; qi += blockDim.x) {
        b.qcount[qi] = 0;
        if (!b.q[qi].trivial) {
            const uint32_t pos = atomicAdd(&s_n, 1u);
            b.items[root_level][pos] = ((uint64_t)qi << 32);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) b.ctr->items[root_level] = s_n;
}

// Items at parent level p are (query, node); the node's <= 64 children at
// level p-1 go on lanes.  A child is culled once k candidates certainly cover
// its whole rectangle (the internal-node test of rknn-inf/inf.h:219-237,
// with certified corner evaluation instead of a 4-corner FP64 test);
// survivors are pushed to level p-1.
__global__ __launch_bounds__(kBlock) void k_batch_internal(DsDev ds, BatchDev b, uint32_t p)
{
    __shared__ CandRec s_rec[kWaves][64];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    const uint32_t stride = gridDim.x * kWaves;
    const uint32_t n_items = min(b.ctr->items[p], b.item_cap[p]);
    const uint32_t cl = p - 1;
    const uint32_t n_child = ds.nlev[cl];
    const float4* __restrict__ cm = ds.mbr[cl];
    uint64_t* __restrict__ next = b.items[cl];
    const uint32_t next_cap = b.item_cap[cl];
    CandRec* wlds = s_rec[wave];
    for (uint32_t it = blockIdx.x * kWaves + wave; it < n_items; it += stride) {
        const uint64_t item = b.items[p][it];
        const uint32_t qi = (uint32_t)(item >> 32), node = (uint32_t)item;
        const QDev Q = b.q[qi];
        const CandRec* __restrict__ rec = b.rec + Q.off;
        const uint32_t c = node * kFanout + lane;
        const bool valid = c < n_child;
        const float4 m = cm[valid ? c : 0];
        const float lx = m.x - Q.qx, ly = m.y - Q.qy, hx = m.z - Q.qx, hy = m.w - Q.qy;
        uint32_t full = 0;
        for (uint32_t c0 = 0; c0 < Q.nc; c0 += 64) {
            const uint32_t n = min(64u, Q.nc - c0);
            if (lane < n) wlds[lane] = rec[c0 + lane];
            __builtin_amdgcn_wave_barrier();
            bool live = true;
            for (uint32_t i0 = 0; i0 < n && live; i0 += 16) {
                const uint32_t i1 = min(n, i0 + 16);
                for (uint32_t i = i0; i < i1; ++i) {
                    const CandRec r = wlds[i];
                    const float t = fmaf(r.gx >= 0.f ? lx : hx, r.gx, fmaf(r.gy >= 0.f ? ly : hy, r.gy, -r.sh));
                    full += (t >= kCertainHi) ? 1u : 0u;
                }
                live = __ballot(valid && full < Q.k) != 0;
            }
            __builtin_amdgcn_wave_barrier();
            if (!live) break;
        }
        const uint64_t surv = __ballot(valid && full < Q.k);
        if (surv) {
            uint32_t pos = 0;
            if (lane == 0) pos = atomicAdd(&b.ctr->items[cl], (uint32_t)__popcll(surv));
            pos = wave_bcast(pos, 0);
            if ((surv >> lane) & 1ull) {
                const uint32_t dst = pos + rank_below(surv);
                if (dst < next_cap)
                    next[dst] = ((uint64_t)qi << 32) | c;
                else
                    b.ctr->overflow = 1;
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_batch_leaf(DsDev ds, BatchDev b)
{
    __shared__ CandRec s_rec[kWaves][64];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    const uint32_t stride = gridDim.x * kWaves;
    const uint32_t n_items = min(b.ctr->items[0], b.item_cap[0]);
    uint32_t rech = 0;
    for (uint32_t it = blockIdx.x * kWaves + wave; it < n_items; it += stride) {
        const uint64_t item = b.items[0][it];
        const uint32_t qi = (uint32_t)(item >> 32), tile = (uint32_t)item;
        const QDev Q = b.q[qi];
        const LeafOut r = leaf_tile(ds, Q.qx, Q.qy, Q.k, Q.nc, b.rec + Q.off, b.orig + Q.off, tile, s_rec[wave]);
        rech += r.rechecks;
        if (r.keep) {
            const uint32_t n = (uint32_t)__popcll(r.keep);
            uint32_t rank0 = 0, pos = 0;
            if (lane == 0) {
                rank0 = atomicAdd(&b.qcount[qi], n);
                pos = atomicAdd(&b.ctr->pool, n);
            }
            rank0 = wave_bcast(rank0, 0);
            pos = wave_bcast(pos, 0);
            if ((r.keep >> lane) & 1ull) {
                const uint32_t rk = rank_below(r.keep);
                const uint32_t dst = pos + rk;
                if (dst < b.pool_cap) {
                    b.pool_uid[dst] = ds.uid[tile * kTile + lane];
                    b.pool_tag[dst] = ((uint64_t)qi << 32) | (rank0 + rk);
                } else {
                    b.ctr->overflow = 1;
                }
            }
        }
    }
    if (lane == 0 && rech) atomicAdd(&b.ctr->rechecks, rech);
}

// Exclusive scan of per-query result counts (one block).
__global__ __launch_bounds__(1024) void k_batch_scan(BatchDev b)
{
    __shared__ uint64_t s_part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (b.nq + 1023) / 1024;
    const uint32_t lo = min(b.nq, t * per), hi = min(b.nq, lo + per);
    uint64_t sum = 0;
    for (uint32_t i = lo; i < hi; ++i) sum += b.qcount[i];
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
        run += b.qcount[i];
    }
    if (t == 1023) b.qoff[b.nq] = s_part[1023];
}

__global__ void k_batch_scatter(BatchDev b)
{
    const uint32_t n = min(b.ctr->pool, b.pool_cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint64_t tag = b.pool_tag[i];
        const uint64_t dst = b.qoff[tag >> 32] + (uint32_t)tag;
        if (dst < b.pool_cap) b.out[dst] = b.pool_uid[i]; // an overflowed batch is discarded by the host
    }
}

// One block per query: bitonic sort of its ids in LDS.  Larger results are
// flagged and sorted by the host when fetched.
__global__ __launch_bounds__(1024) void k_batch_sort(BatchDev b)
{
    __shared__ int32_t s[kSortCap];
    for (uint32_t qi = blockIdx.x; qi < b.nq; qi += gridDim.x) {
        const uint32_t n = b.qcount[qi];
        if (n <= 1 || b.qoff[qi] + n > b.pool_cap) continue; // overflowed batches are discarded
        if (n > kSortCap) {
            if (threadIdx.x == 0) atomicAdd(&b.ctr->sort_overflow, 1u);
            continue;
        }
        uint32_t np = 1;
        while (np < n) np <<= 1;
        int32_t* seg = b.out + b.qoff[qi];
        for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) s[i] = i < n ? seg[i] : INT32_MAX;
        __syncthreads();
        for (uint32_t size = 2; size <= np; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t i = threadIdx.x; i < np / 2; i += blockDim.x) {
                    const uint32_t lo = 2 * i - (i & (stride - 1));
                    const uint32_t hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const int32_t a = s[lo], c = s[hi];
                    if ((a > c) == up) {
                        s[lo] = c;
                        s[hi] = a;
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) seg[i] = s[i];
        __syncthreads();
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

void launch_morton_keys(const float* ux, const float* uy, uint32_t n, float minx, float miny, float sx, float sy,
                        uint32_t* keys, uint32_t* idx, hipStream_t st)
{
    k_morton<<<grid_for(n, 256, 4096), 256, 0, st>>>(ux, uy, n, minx, miny, sx, sy, keys, idx);
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

void launch_prep_records(const QDev* q, uint32_t nq, const uint32_t* cand, const float* fx, const float* fy,
                         const double2* wbound, CandRec* rec, float2* orig, uint64_t total, hipStream_t st)
{
    k_prep_records<<<nq < 4096 ? (nq ? nq : 1) : 4096, 256, 0, st>>>(q, nq, cand, fx, fy, wbound, rec, orig, total);
}

void launch_batch_init(const BatchDev& b, uint32_t root_level, hipStream_t st)
{
    k_batch_init<<<1, 1024, 0, st>>>(b, root_level);
}

void launch_batch_internal(const DsDev& ds, const BatchDev& b, uint32_t parent_level, int grid, hipStream_t st)
{
    k_batch_internal<<<grid, kBlock, 0, st>>>(ds, b, parent_level);
}

void launch_batch_leaf(const DsDev& ds, const BatchDev& b, int grid, hipStream_t st)
{
    k_batch_leaf<<<grid, kBlock, 0, st>>>(ds, b);
}

void launch_batch_finish(const BatchDev& b, int grid, hipStream_t st)
{
    k_batch_scan<<<1, 1024, 0, st>>>(b);
    k_batch_scatter<<<grid, 256, 0, st>>>(b);
    k_batch_sort<<<b.nq < 65535 ? (b.nq ? b.nq : 1) : 65535, 1024, 0, st>>>(b);
}

} // namespace rk
