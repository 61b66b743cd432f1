// rk_api.hip -- C ABI of include/rknn.h: contexts, resident datasets and the
// plain, brute-force, filtered and count-all query forms.  The batched form
// is in rk_batch.hip and the serving loop in rk_serve.cpp.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rk_internal.hpp"
#include "rknn.h"

namespace rk {

rk_status fail(rk_ctx* c, rk_status s, const std::string& msg)
{
    if (c) c->err = msg;
    return s;
}

rk_status fail_hip(rk_ctx* c, hipError_t e, const char* what, int line)
{
    std::string m = std::string(what) + " failed at rk line " + std::to_string(line) + ": " + hipGetErrorString(e);
    return fail(c, e == hipErrorOutOfMemory ? RK_ERR_OUT_OF_MEMORY : RK_ERR_HIP, m);
}

rk_status clean_candidates(rk_dataset* ds, uint32_t q, const uint32_t* cand, size_t nc, std::vector<uint32_t>& out)
{
    out.clear();
    if (nc && !cand) return fail(ds->ctx, RK_ERR_INVALID_ARG, "candidate list is NULL with nc > 0");
    out.reserve(nc);
    const float qx = ds->hfx[q], qy = ds->hfy[q];
    for (size_t i = 0; i < nc; ++i) {
        const uint32_t f = cand[i];
        if (f >= ds->nf) return fail(ds->ctx, RK_ERR_INVALID_ARG, "candidate index out of range");
        if (f == q || (ds->hfx[f] == qx && ds->hfy[f] == qy)) continue;
        out.push_back(f);
    }
    return RK_OK;
}

} // namespace rk

using namespace rk;

// ------------------------------------------------- allocation-free scratch
//
// rk_query_plain_noalloc's device memory beyond its inputs and outputs: the
// query's candidate indices, their records, the plain kernel's counters and
// the staging of the result ids.  One array per device, held until the
// process ends; each device's copy lives at its own address, which
// hipGetSymbolAddress reports (once per device, at rk_init) and which host
// code uses in place of the array's name.
namespace {

constexpr uint32_t kNaMaxCand = RK_NOALLOC_MAX_CANDIDATES;
constexpr uint32_t kNaMaxIds = RK_NOALLOC_MAX_IDS;
constexpr size_t kNaCandOff = 0;
constexpr size_t kNaRecOff = kNaCandOff + kNaMaxCand * sizeof(uint32_t);
constexpr size_t kNaOrigOff = kNaRecOff + kNaMaxCand * sizeof(CandRec);
constexpr size_t kNaCtrOff = kNaOrigOff + kNaMaxCand * sizeof(float2);
constexpr size_t kNaIdsOff = kNaCtrOff + 256 * sizeof(uint32_t);
constexpr size_t kNaBytes = kNaIdsOff + (size_t)kNaMaxIds * sizeof(int32_t);
static_assert(kPlainCounters <= 256, "counter slots");
static_assert(kNaRecOff % 16 == 0 && kNaOrigOff % 16 == 0 && kNaIdsOff % 16 == 0, "scratch alignment");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;                            // guards g_na_ptr
unsigned char* g_na_ptr[kNaMaxDevices] = {};   // per device: the scratch array's device address
std::mutex g_na_call[kNaMaxDevices];           // per device: one allocation-free call at a time

// With `device` current: the device address of its scratch array, looked
// up once per device.
hipError_t scratch_address(int device, unsigned char** out)
{
    *out = nullptr;
    if (device < 0 || device >= kNaMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_na_ptr[device] = static_cast<unsigned char*>(p);
    }
    *out = g_na_ptr[device];
    return hipSuccess;
}

} // namespace

namespace {

bool finite_bounded(const float* v, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i]) || std::fabs(v[i]) > 1e18f) return false;
    return true;
}

rk_status check_query(rk_dataset* ds, uint32_t q, uint32_t k)
{
    if (!ds) return RK_ERR_INVALID_ARG;
    if (q >= ds->nf) return fail(ds->ctx, RK_ERR_INVALID_ARG, "query index q must be < number of facilities");
    if (k == 0) return fail(ds->ctx, RK_ERR_INVALID_ARG, "k must be a positive integer");
    return RK_OK;
}

// Deliver a sorted id list to the caller's buffer.
rk_status deliver(rk_dataset* ds, const int32_t* ids, size_t n, int32_t* out, size_t cap, size_t* n_out)
{
    *n_out = n;
    if (n > cap) return fail(ds->ctx, RK_ERR_CAPACITY, "output buffer too small; *n_out holds the size needed");
    if (n) std::memcpy(out, ids, n * sizeof(int32_t));
    return RK_OK;
}

// Host side of the plain and count-all forms: binary64 records of the
// cleaned candidate list into pinned memory, then one async upload.
rk_status upload_records(rk_dataset* ds, uint32_t q, const std::vector<uint32_t>& cl)
{
    rk_ctx* c = ds->ctx;
    const size_t n = cl.size();
    RK_HIP(c, ds->h_rec.reserve(n));
    RK_HIP(c, ds->h_orig.reserve(n));
    RK_HIP(c, ds->rec.reserve(n));
    RK_HIP(c, ds->orig.reserve(n));
    const float qx = ds->hfx[q], qy = ds->hfy[q];
    double wx, wy;
    axis_bounds(ds->ubox[0], ds->ubox[1], ds->ubox[2], ds->ubox[3], qx, qy, &wx, &wy);
    for (size_t j = 0; j < n; ++j) {
        const uint32_t f = cl[j];
        ds->h_rec.p[j] = make_record(qx, qy, ds->hfx[f], ds->hfy[f], wx, wy);
        ds->h_orig.p[j] = make_float2(ds->hfx[f], ds->hfy[f]);
    }
    if (n) {
        RK_HIP(c, hipMemcpyAsync(ds->rec.p, ds->h_rec.p, n * sizeof(CandRec), hipMemcpyHostToDevice, c->stream));
        RK_HIP(c, hipMemcpyAsync(ds->orig.p, ds->h_orig.p, n * sizeof(float2), hipMemcpyHostToDevice, c->stream));
    }
    return RK_OK;
}

// Count-all kernel class: device and candidates (power of two).
std::string count_key(rk_dataset* ds, size_t nc)
{
    uint64_t p = 1, u = 1;
    while (p < nc && p < (1ull << 30)) p <<= 1;
    while (u < ds->nu && u < (1ull << 40)) u <<= 1;
    return device_arch(ds->ctx) + "/count/u" + std::to_string(u) + "/c" + std::to_string(p);
}

// RK_COUNT_KERNEL=valu4|valu6|valu8 forces a variant; otherwise the
// autotuner's choice for the class, or -1 when the class has not been timed
// yet.
int count_variant(rk_dataset* ds, size_t nc)
{
    if (const char* e = std::getenv("RK_COUNT_KERNEL")) {
        for (int v = 0; v < kCountVariants; ++v)
            if (std::strcmp(e, count_variant_name(v)) == 0) return v;
    }
    const int v = tuned_choice(count_key(ds, nc));
    return v < kCountVariants ? v : -1;
}

QDev make_qdev(rk_dataset* ds, uint32_t q, uint32_t k, size_t nc)
{
    QDev d{};
    d.qx = ds->hfx[q];
    d.qy = ds->hfy[q];
    d.k = k;
    d.nc = (uint32_t)nc;
    d.off = 0;
    d.trivial = k > nc ? 1u : 0u;
    return d;
}

rk_status plain_core(rk_dataset* ds, uint32_t q, uint32_t k, const std::vector<uint32_t>& cl, int32_t* out_ids,
                     size_t cap, size_t* n_out, rk_timing* t, double t_start)
{
    rk_ctx* c = ds->ctx;
    if (ds->nu == 0) {
        *n_out = 0;
        return RK_OK;
    }
    if (k > cl.size()) { // fewer candidates than k: every user qualifies
        rk_status s = deliver(ds, ds->all_ids_sorted.data(), ds->all_ids_sorted.size(), out_ids, cap, n_out);
        if (t) t->total_s = now_s() - t_start;
        return s;
    }
    const double t0 = now_s();
    rk_status s = upload_records(ds, q, cl);
    if (s != RK_OK) return s;
    RK_HIP(c, ds->ctr.reserve(kPlainCounters));
    RK_HIP(c, ds->h_ctr.reserve(kPlainCounters));
    RK_HIP(c, hipMemsetAsync(ds->ctr.p, 0, kPlainCounters * sizeof(uint32_t), c->stream));
    const double t1 = now_s();
    QDev qd = make_qdev(ds, q, k, cl.size());
    RK_HIP(c, hipEventRecord(c->ev0, c->stream));
    launch_plain(ds->view, qd, ds->rec.p, ds->orig.p, ds->out.p, ds->n_pad, ds->ctr.p, device_grid(c->n_cu, 32),
                 c->stream);
    RK_HIP(c, hipGetLastError());
    RK_HIP(c, hipEventRecord(c->ev1, c->stream));
    RK_HIP(c, hipMemcpyAsync(ds->h_ctr.p, ds->ctr.p, kPlainCounters * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->stream));
    RK_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    RK_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    const uint32_t n = ds->h_ctr.p[0];
    uint64_t fp64 = 0;
    for (int j = 1; j < kPlainCounters; ++j) fp64 += ds->h_ctr.p[j];
    if (n > ds->n_pad) return fail(c, RK_ERR_INTERNAL, "result count exceeds the user count");
    RK_HIP(c, ds->h_out.reserve(n));
    if (n) {
        RK_HIP(c, hipMemcpyAsync(ds->h_out.p, ds->out.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        RK_HIP(c, hipStreamSynchronize(c->stream));
    }
    std::sort(ds->h_out.p, ds->h_out.p + n);
    s = deliver(ds, ds->h_out.p, n, out_ids, cap, n_out);
    if (t) {
        t->upload_s = t1 - t0;
        t->kernel_s = ms * 1e-3;
        t->n_candidates = cl.size();
        t->n_fp64_rechecks = fp64;
        t->total_s = now_s() - t_start;
        t->download_s = t->total_s - (t1 - t_start) - t->kernel_s;
    }
    return s;
}

} // namespace

extern "C" {

const char* rk_status_string(rk_status s)
{
    switch (s) {
    case RK_OK: return "ok";
    case RK_ERR_NO_DEVICE: return "no HIP device available";
    case RK_ERR_INVALID_ARG: return "invalid argument";
    case RK_ERR_OUT_OF_MEMORY: return "device out of memory";
    case RK_ERR_CAPACITY: return "output buffer too small";
    case RK_ERR_HIP: return "HIP runtime error";
    case RK_ERR_IO: return "input/output error";
    case RK_ERR_ADMISSION: return "request rejected by memory admission";
    case RK_ERR_NOT_READY: return "request still in flight";
    case RK_ERR_INTERNAL: return "internal error";
    }
    return "unknown status";
}

const char* rk_last_error(const rk_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

int rk_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

rk_status rk_init(int device, rk_ctx** out)
{
    if (!out) return RK_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RK_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return RK_ERR_NO_DEVICE;
    rk_ctx* c = new rk_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
        delete c;
        return RK_ERR_HIP;
    }
    if (c->n_cu <= 0) c->n_cu = 1;
    if (scratch_address(device, &c->scratch) != hipSuccess) {
        (void)hipEventDestroy(c->ev0);
        (void)hipEventDestroy(c->ev1);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return RK_ERR_HIP;
    }
    *out = c;
    return RK_OK;
}

rk_status rk_shutdown(rk_ctx* c)
{
    if (!c) return RK_ERR_INVALID_ARG;
    {
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        (void)hipSetDevice(c->device);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->ev1) (void)hipEventDestroy(c->ev1);
        for (hipEvent_t e : c->ring_ev)
            if (e) (void)hipEventDestroy(e);
        c->gang.reset();
        if (c->stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return RK_OK;
}

rk_status rk_dataset_create(rk_ctx* c, const float* ux, const float* uy, const int32_t* uid, size_t nu,
                            const float* fx, const float* fy, const int32_t* fid, size_t nf, rk_dataset** out)
{
    if (!c || !out) return RK_ERR_INVALID_ARG;
    *out = nullptr;
    if ((nu && (!ux || !uy || !uid)) || (nf && (!fx || !fy || !fid)))
        return fail(c, RK_ERR_INVALID_ARG, "NULL coordinate or id array");
    if (nu > 0x7fffff00ull || nf > 0x7fffff00ull) return fail(c, RK_ERR_INVALID_ARG, "too many points");
    if (!finite_bounded(ux, nu) || !finite_bounded(uy, nu) || !finite_bounded(fx, nf) || !finite_bounded(fy, nf))
        return fail(c, RK_ERR_INVALID_ARG, "coordinates must be finite with magnitude <= 1e18");
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    std::unique_ptr<rk_dataset> ds(new rk_dataset());
    ds->ctx = c;
    ds->nu = nu;
    ds->nf = nf;
    ds->hfx.assign(fx, fx + nf);
    ds->hfy.assign(fy, fy + nf);
    ds->hfid.assign(fid, fid + nf);
    if (nu) {
        double b[4] = {ux[0], uy[0], ux[0], uy[0]};
        for (size_t i = 1; i < nu; ++i) {
            b[0] = std::min(b[0], (double)ux[i]);
            b[1] = std::min(b[1], (double)uy[i]);
            b[2] = std::max(b[2], (double)ux[i]);
            b[3] = std::max(b[3], (double)uy[i]);
        }
        std::memcpy(ds->ubox, b, sizeof b);
        ds->all_ids_sorted.assign(uid, uid + nu);
        std::sort(ds->all_ids_sorted.begin(), ds->all_ids_sorted.end());
        if (ds->all_ids_sorted.front() >= 0) { // sort keys need only the bits below the largest id
            uint32_t bits = 1;
            while (bits < 31 && ((uint32_t)ds->all_ids_sorted.back() >> bits) != 0) ++bits;
            ds->id_bits = bits;
        }
    }
    hipStream_t st = c->stream;
    if (nf) {
        RK_HIP(c, ds->fx.reserve(nf));
        RK_HIP(c, ds->fy.reserve(nf));
        RK_HIP(c, hipMemcpyAsync(ds->fx.p, fx, nf * sizeof(float), hipMemcpyHostToDevice, st));
        RK_HIP(c, hipMemcpyAsync(ds->fy.p, fy, nf * sizeof(float), hipMemcpyHostToDevice, st));
    }
    ds->index.reset(new FacilityIndex(fx, fy, nf, ds->ubox));
    {
        const rk_status gs = build_facility_grid(ds.get());
        if (gs != RK_OK) return gs;
    }

    ds->n_tiles = (uint32_t)((nu + kTile - 1) / kTile);
    ds->n_pad = ds->n_tiles * kTile;
    if (nu) {
        const uint32_t n = (uint32_t)nu;
        DevBuf<float> tx, ty;
        DevBuf<int32_t> tid;
        DevBuf<uint32_t> keys, keys2, idx;
        RK_HIP(c, tx.reserve(n));
        RK_HIP(c, ty.reserve(n));
        RK_HIP(c, tid.reserve(n));
        RK_HIP(c, keys.reserve(n));
        RK_HIP(c, keys2.reserve(n));
        RK_HIP(c, idx.reserve(n));
        RK_HIP(c, ds->perm.reserve(n));
        RK_HIP(c, hipMemcpyAsync(tx.p, ux, n * sizeof(float), hipMemcpyHostToDevice, st));
        RK_HIP(c, hipMemcpyAsync(ty.p, uy, n * sizeof(float), hipMemcpyHostToDevice, st));
        RK_HIP(c, hipMemcpyAsync(tid.p, uid, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        const double w = ds->ubox[2] - ds->ubox[0], h = ds->ubox[3] - ds->ubox[1];
        const float sx = w > 0 ? (float)(65535.0 / w) : 0.f, sy = h > 0 ? (float)(65535.0 / h) : 0.f;
        launch_curve_keys(tx.p, ty.p, n, (float)ds->ubox[0], (float)ds->ubox[1], sx, sy, keys.p, idx.p, st);
        RK_HIP(c, hipGetLastError());
        size_t tmp_bytes = 0;
        RK_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys.p, keys2.p, idx.p, ds->perm.p, (int)n,
                                                     0, 32, st));
        DevBuf<unsigned char> tmp;
        RK_HIP(c, tmp.reserve(tmp_bytes));
        RK_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, keys.p, keys2.p, idx.p, ds->perm.p, (int)n, 0,
                                                     32, st));
        RK_HIP(c, ds->ux.reserve(ds->n_pad));
        RK_HIP(c, ds->uy.reserve(ds->n_pad));
        RK_HIP(c, ds->uid.reserve(ds->n_pad));
        launch_gather_users(ds->perm.p, n, ds->n_pad, tx.p, ty.p, tid.p, ds->ux.p, ds->uy.p, ds->uid.p, st);
        RK_HIP(c, hipGetLastError());
        // bounding-rectangle tree: tiles, then groups of 64 until <= 64 nodes
        uint32_t cnt = ds->n_tiles;
        int lvl = 0;
        RK_HIP(c, ds->mbr[0].reserve(cnt));
        launch_tile_mbr(ds->ux.p, ds->uy.p, cnt, ds->mbr[0].p, st);
        RK_HIP(c, hipGetLastError());
        ds->view.nlev[0] = cnt;
        while (cnt > (uint32_t)kFanout) {
            if (lvl + 1 >= kMaxLevels) return fail(c, RK_ERR_INVALID_ARG, "dataset too large for the tree depth");
            const uint32_t up = (cnt + kFanout - 1) / kFanout;
            RK_HIP(c, ds->mbr[lvl + 1].reserve(up));
            launch_level_mbr(ds->mbr[lvl].p, cnt, ds->mbr[lvl + 1].p, up, st);
            RK_HIP(c, hipGetLastError());
            ++lvl;
            cnt = up;
            ds->view.nlev[lvl] = cnt;
        }
        ds->view.n_levels = (uint32_t)lvl + 1;
        // joint rectangles of the count-all kernels' 4- and 8-tile groups
        RK_HIP(c, ds->gmbr4.reserve((ds->n_tiles + 3) / 4));
        RK_HIP(c, ds->gmbr6.reserve((ds->n_tiles + 5) / 6));
        RK_HIP(c, ds->gmbr8.reserve((ds->n_tiles + 7) / 8));
        launch_group_mbr(ds->mbr[0].p, ds->n_tiles, 4, ds->gmbr4.p, st);
        RK_HIP(c, hipGetLastError());
        launch_group_mbr(ds->mbr[0].p, ds->n_tiles, 6, ds->gmbr6.p, st);
        RK_HIP(c, hipGetLastError());
        launch_group_mbr(ds->mbr[0].p, ds->n_tiles, 8, ds->gmbr8.p, st);
        RK_HIP(c, hipGetLastError());
        RK_HIP(c, hipStreamSynchronize(st)); // temporaries are freed on return
        RK_HIP(c, ds->out.reserve(ds->n_pad));
    }
    ds->view.ux = ds->ux.p;
    ds->view.uy = ds->uy.p;
    ds->view.uid = ds->uid.p;
    for (int l = 0; l < kMaxLevels; ++l) ds->view.mbr[l] = ds->mbr[l].p;
    ds->view.gmbr4 = ds->gmbr4.p;
    ds->view.gmbr6 = ds->gmbr6.p;
    ds->view.gmbr8 = ds->gmbr8.p;
    ds->view.nu = (uint32_t)nu;
    ds->view.n_tiles = ds->n_tiles;
    RK_HIP(c, hipStreamSynchronize(st));
    *out = ds.release();
    return RK_OK;
}

rk_status rk_dataset_destroy(rk_dataset* ds)
{
    if (!ds) return RK_ERR_INVALID_ARG;
    rk_ctx* c = ds->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete ds;
    return RK_OK;
}

rk_status rk_dataset_info(const rk_dataset* ds, uint64_t* nu, uint64_t* nf, uint64_t* device_bytes)
{
    if (!ds) return RK_ERR_INVALID_ARG;
    if (nu) *nu = ds->nu;
    if (nf) *nf = ds->nf;
    if (device_bytes) {
        uint64_t b = ds->ux.bytes() + ds->uy.bytes() + ds->uid.bytes() + ds->perm.bytes() + ds->fx.bytes() +
                     ds->fy.bytes() + ds->out.bytes();
        for (int l = 0; l < kMaxLevels; ++l) b += ds->mbr[l].bytes();
        b += ds->gmbr4.bytes() + ds->gmbr6.bytes() + ds->gmbr8.bytes();
        b += ds->g_start.bytes() + ds->g_fid.bytes() + ds->g_fx.bytes() + ds->g_fy.bytes() + ds->g_dir.bytes();
        *device_bytes = b;
    }
    return RK_OK;
}

rk_status rk_filter(rk_dataset* ds, uint32_t q, uint32_t k, uint32_t* cand, size_t cap, size_t* nc)
{
    rk_status s = check_query(ds, q, k);
    if (s != RK_OK) return s;
    if (!nc || (cap && !cand)) return fail(ds->ctx, RK_ERR_INVALID_ARG, "NULL output");
    std::vector<uint32_t> v;
    ds->index->filter(q, k, v);
    *nc = v.size();
    if (v.size() > cap) return fail(ds->ctx, RK_ERR_CAPACITY, "candidate buffer too small; *nc holds the size needed");
    if (!v.empty()) std::memcpy(cand, v.data(), v.size() * sizeof(uint32_t));
    return RK_OK;
}

rk_status rk_query_plain(rk_dataset* ds, uint32_t q, uint32_t k, const uint32_t* cand, size_t nc, int32_t* out_ids,
                         size_t cap, size_t* n_out, rk_timing* t)
{
    const double t0 = now_s();
    rk_status s = check_query(ds, q, k);
    if (s != RK_OK) return s;
    if (!n_out || (cap && !out_ids)) return fail(ds->ctx, RK_ERR_INVALID_ARG, "NULL output");
    std::vector<uint32_t> cl;
    s = clean_candidates(ds, q, cand, nc, cl);
    if (s != RK_OK) return s;
    if (t) *t = rk_timing{};
    std::lock_guard<std::recursive_mutex> lk(ds->ctx->mu);
    RK_HIP(ds->ctx, hipSetDevice(ds->ctx->device));
    return plain_core(ds, q, k, cl, out_ids, cap, n_out, t, t0);
}

rk_status rk_query_bruteforce(rk_dataset* ds, uint32_t q, uint32_t k, int32_t* out_ids, size_t cap, size_t* n_out,
                              rk_timing* t)
{
    const double t0 = now_s();
    rk_status s = check_query(ds, q, k);
    if (s != RK_OK) return s;
    if (!n_out || (cap && !out_ids)) return fail(ds->ctx, RK_ERR_INVALID_ARG, "NULL output");
    std::vector<uint32_t> all(ds->nf), cl;
    for (uint64_t i = 0; i < ds->nf; ++i) all[i] = (uint32_t)i;
    s = clean_candidates(ds, q, all.data(), all.size(), cl);
    if (s != RK_OK) return s;
    if (t) *t = rk_timing{};
    std::lock_guard<std::recursive_mutex> lk(ds->ctx->mu);
    RK_HIP(ds->ctx, hipSetDevice(ds->ctx->device));
    return plain_core(ds, q, k, cl, out_ids, cap, n_out, t, t0);
}

rk_status rk_query(rk_dataset* ds, uint32_t q, uint32_t k, int32_t* out_ids, size_t cap, size_t* n_out, rk_timing* t)
{
    const double t0 = now_s();
    rk_status s = check_query(ds, q, k);
    if (s != RK_OK) return s;
    if (!n_out || (cap && !out_ids)) return fail(ds->ctx, RK_ERR_INVALID_ARG, "NULL output");
    std::vector<uint32_t> cand, cl;
    ds->index->filter(q, k, cand);
    const double t1 = now_s();
    s = clean_candidates(ds, q, cand.data(), cand.size(), cl);
    if (s != RK_OK) return s;
    if (t) *t = rk_timing{};
    std::lock_guard<std::recursive_mutex> lk(ds->ctx->mu);
    RK_HIP(ds->ctx, hipSetDevice(ds->ctx->device));
    s = plain_core(ds, q, k, cl, out_ids, cap, n_out, t, t0);
    if (t) t->filter_s = t1 - t0;
    return s;
}

rk_status rk_query_plain_noalloc(rk_dataset* ds, uint32_t q, uint32_t k, const uint32_t* cand, size_t nc,
                                 int32_t* out_ids, size_t cap, size_t* n_out)
{
    rk_status s = check_query(ds, q, k);
    if (s != RK_OK) return s;
    rk_ctx* c = ds->ctx;
    if (!n_out || (cap && !out_ids)) return fail(c, RK_ERR_INVALID_ARG, "NULL output");
    if (nc && !cand) return fail(c, RK_ERR_INVALID_ARG, "candidate list is NULL with nc > 0");
    if (nc > kNaMaxCand)
        return fail(c, RK_ERR_INVALID_ARG,
                    "the allocation-free form takes at most RK_NOALLOC_MAX_CANDIDATES (4096) candidates");
    if (!c->scratch) return fail(c, RK_ERR_INTERNAL, "no scratch address for this device");
    // the candidate list cleaned as clean_candidates does, counted in place
    const float qx = ds->hfx[q], qy = ds->hfy[q];
    size_t live = 0;
    for (size_t i = 0; i < nc; ++i) {
        const uint32_t f = cand[i];
        if (f >= ds->nf) return fail(c, RK_ERR_INVALID_ARG, "candidate index out of range");
        live += (f == q || (ds->hfx[f] == qx && ds->hfy[f] == qy)) ? 0 : 1;
    }
    *n_out = 0;
    if (ds->nu == 0) return RK_OK;
    if (k > live) // fewer candidates than k: every user qualifies
        return deliver(ds, ds->all_ids_sorted.data(), ds->all_ids_sorted.size(), out_ids, cap, n_out);
    double wx = 0, wy = 0;
    axis_bounds(ds->ubox[0], ds->ubox[1], ds->ubox[2], ds->ubox[3], qx, qy, &wx, &wy);
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    std::lock_guard<std::mutex> one(g_na_call[c->device]); // contexts sharing the device share the array
    RK_HIP(c, hipSetDevice(c->device));
    unsigned char* base = c->scratch;
    uint32_t* d_cand = reinterpret_cast<uint32_t*>(base + kNaCandOff);
    CandRec* d_rec = reinterpret_cast<CandRec*>(base + kNaRecOff);
    float2* d_orig = reinterpret_cast<float2*>(base + kNaOrigOff);
    uint32_t* d_ctr = reinterpret_cast<uint32_t*>(base + kNaCtrOff);
    int32_t* d_ids = reinterpret_cast<int32_t*>(base + kNaIdsOff);
    hipStream_t st = c->stream;
    if (nc) RK_HIP(c, hipMemcpyAsync(d_cand, cand, nc * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    // records of every listed candidate (dead ones for q and its position)
    // and the counters cleared by the same launch
    launch_build_records(d_cand, (uint32_t)nc, q, qx, qy, wx, wy, ds->fx.p, ds->fy.p, d_rec, d_orig, d_ctr,
                         kPlainCounters, st);
    RK_HIP(c, hipGetLastError());
    QDev qd = make_qdev(ds, q, k, nc);
    launch_plain(ds->view, qd, d_rec, d_orig, d_ids, kNaMaxIds, d_ctr, device_grid(c->n_cu, 32), st);
    RK_HIP(c, hipGetLastError());
    uint32_t n = 0;
    RK_HIP(c, hipMemcpyAsync(&n, d_ctr, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    RK_HIP(c, hipStreamSynchronize(st));
    *n_out = n;
    if (n > kNaMaxIds)
        return fail(c, RK_ERR_CAPACITY,
                    "the result exceeds the allocation-free form's staging of RK_NOALLOC_MAX_IDS (1048576) ids; "
                    "*n_out holds its size");
    if (n > cap) return fail(c, RK_ERR_CAPACITY, "output buffer too small; *n_out holds the size needed");
    if (n) {
        RK_HIP(c, hipMemcpyAsync(out_ids, d_ids, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        RK_HIP(c, hipStreamSynchronize(st));
    }
    std::sort(out_ids, out_ids + n); // in place: no allocation
    return RK_OK;
}

} // extern "C"

namespace {

// dst is device memory of the given device: kernels running there may write
// it (results stay off PCIe).  Device memory of another device is reached by
// a copy instead (download).
bool device_dst(const void* dst, int device)
{
    hipPointerAttribute_t a{};
    const hipError_t pe = hipPointerGetAttributes(&a, dst);
    (void)hipGetLastError(); // a pageable pointer may leave an error behind
    return pe == hipSuccess && a.type == hipMemoryTypeDevice && a.device == device;
}

// Copies bytes from device memory to dst, in stream order on the context's
// stream, and waits for them.  Device memory at dst takes a device copy; a
// pinned destination (hipHostMalloc) one DMA at the PCIe rate; pageable
// memory, which a plain hipMemcpy fills at about a third of it, goes
// through two pinned chunks: the DMA of one chunk runs while the context's
// host threads copy the other into place.
rk_status download(rk_ctx* c, void* dst, const void* src, size_t bytes)
{
    constexpr size_t kChunk = size_t(8) << 20;
    hipPointerAttribute_t a{};
    const hipError_t pe = hipPointerGetAttributes(&a, dst);
    (void)hipGetLastError(); // a pageable pointer may leave an error behind
    const bool pinned = pe == hipSuccess && a.type == hipMemoryTypeHost;
    if (pe == hipSuccess && (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
                             a.type == hipMemoryTypeUnified)) {
        RK_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, c->stream));
        RK_HIP(c, hipStreamSynchronize(c->stream));
        return RK_OK;
    }
    if (pinned || bytes < 2 * kChunk) {
        RK_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        RK_HIP(c, hipStreamSynchronize(c->stream));
        return RK_OK;
    }
    RK_HIP(c, c->ring.reserve(2 * kChunk));
    for (hipEvent_t& e : c->ring_ev)
        if (!e) RK_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    Gang& gang = host_gang(c);
    const size_t n = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](size_t i) -> rk_status {
        const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
        RK_HIP(c, hipMemcpyAsync(c->ring.p + (i & 1) * kChunk, (const unsigned char*)src + off, len,
                                 hipMemcpyDeviceToHost, c->stream));
        RK_HIP(c, hipEventRecord(c->ring_ev[i & 1], c->stream));
        return RK_OK;
    };
    rk_status s = issue(0);
    if (s == RK_OK) s = issue(1);
    for (size_t i = 0; i < n && s == RK_OK; ++i) {
        RK_HIP(c, hipEventSynchronize(c->ring_ev[i & 1]));
        const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
        const unsigned char* from = c->ring.p + (i & 1) * kChunk;
        unsigned char* to = (unsigned char*)dst + off;
        const unsigned t = gang.size();
        gang.run([&](unsigned k) {
            const size_t lo = len * k / t, hi = len * (k + 1) / t;
            std::memcpy(to + lo, from + lo, hi - lo);
        });
        if (i + 2 < n) s = issue(i + 2);
    }
    if (s != RK_OK) (void)hipStreamSynchronize(c->stream);
    return s;
}

// rk_count_all and rk_count_all_sorted: counts in the caller's user order
// (a scatter through perm after the count) or in the dataset's sorted order
// (32-bit counts straight from the count kernel).
rk_status count_all_impl(rk_dataset* ds, uint32_t q, const uint32_t* cand, size_t nc, uint32_t* counts,
                         rk_timing* t, bool sorted_out);
rk_status count_many_impl(rk_dataset* ds, const rk_query_desc* qs, size_t nq, uint32_t* counts, size_t stride,
                          bool sorted_out, rk_timing* t);

} // namespace

extern "C" {

rk_status rk_count_all(rk_dataset* ds, uint32_t q, const uint32_t* cand, size_t nc, uint32_t* counts, rk_timing* t)
{
    return count_all_impl(ds, q, cand, nc, counts, t, false);
}

rk_status rk_count_all_sorted(rk_dataset* ds, uint32_t q, const uint32_t* cand, size_t nc, uint32_t* counts,
                              rk_timing* t)
{
    return count_all_impl(ds, q, cand, nc, counts, t, true);
}

rk_status rk_count_all_many(rk_dataset* ds, const rk_query_desc* qs, size_t nq, uint32_t* counts, size_t stride,
                            int sorted_order, rk_timing* t)
{
    return count_many_impl(ds, qs, nq, counts, stride, sorted_order != 0, t);
}

rk_status rk_dataset_order(const rk_dataset* ds, uint32_t* order)
{
    if (!ds) return RK_ERR_INVALID_ARG;
    if (ds->nu && !order) return fail(ds->ctx, RK_ERR_INVALID_ARG, "NULL output");
    if (!ds->nu) return RK_OK;
    rk_ctx* c = ds->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    RK_HIP(c, hipMemcpyAsync(order, ds->perm.p, ds->nu * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    RK_HIP(c, hipStreamSynchronize(c->stream));
    return RK_OK;
}

} // extern "C"

namespace {

// The count kernel for qd's class: RK_COUNT_KERNEL, the autotuner's stored
// choice, or -- on the first query of a class not yet timed -- every variant
// timed twice on this query (records at rec, counts into counts) and the
// fastest kept.  The binary64 counters are cleared again after a timing.
rk_status pick_count_variant(rk_dataset* ds, const QDev& qd, const CandRec* rec, const float2* orig, void* counts,
                             bool wide, int* out)
{
    rk_ctx* c = ds->ctx;
    int variant = count_variant(ds, qd.nc);
    if (variant >= 0) {
        *out = variant;
        return RK_OK;
    }
    float best = 0.f;
    for (int v = 0; v < kCountVariants; ++v) {
        float ms = 0.f;
        for (int r = 0; r < 2; ++r) {
            RK_HIP(c, hipEventRecord(c->ev0, c->stream));
            launch_count_all(ds->view, qd, rec, orig, counts, ds->ctr.p + 1, v, c->stream, wide);
            RK_HIP(c, hipGetLastError());
            RK_HIP(c, hipEventRecord(c->ev1, c->stream));
            RK_HIP(c, hipEventSynchronize(c->ev1));
            RK_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        }
        if (v == 0 || ms < best) {
            best = ms;
            variant = v;
        }
    }
    tuned_store(count_key(ds, qd.nc), variant, count_variant_name(variant));
    RK_HIP(c, hipMemsetAsync(ds->ctr.p, 0, kPlainCounters * sizeof(uint32_t), c->stream));
    *out = variant;
    return RK_OK;
}

// Device bytes of one round of rk_count_all_many's count slices (a single
// query's slice may exceed it: then the round holds that query alone).
constexpr size_t kCountManyBytes = size_t(256) << 20;

rk_status count_many_impl(rk_dataset* ds, const rk_query_desc* qs, size_t nq, uint32_t* counts, size_t stride,
                          bool sorted_out, rk_timing* t)
{
    const double t0 = now_s();
    if (!ds) return RK_ERR_INVALID_ARG;
    if (nq && !qs) return fail(ds->ctx, RK_ERR_INVALID_ARG, "NULL query array");
    if (nq && ds->nu && !counts) return fail(ds->ctx, RK_ERR_INVALID_ARG, "NULL output");
    if (nq && stride < ds->nu) return fail(ds->ctx, RK_ERR_INVALID_ARG, "stride must be at least the number of users");
    if (t) *t = rk_timing{};
    // every query's cleaned list; the host filter (at the query's k) for a
    // query without one
    std::vector<std::vector<uint32_t>> cls(nq);
    double filter_s = 0;
    for (size_t i = 0; i < nq; ++i) {
        const rk_query_desc& d = qs[i];
        rk_status s = check_query(ds, d.q, d.cand ? 1u : d.k);
        if (s != RK_OK) return s;
        if (d.cand) {
            s = clean_candidates(ds, d.q, d.cand, d.nc, cls[i]);
        } else {
            std::vector<uint32_t> v;
            const double f0 = now_s();
            ds->index->filter(d.q, d.k, v);
            filter_s += now_s() - f0;
            s = clean_candidates(ds, d.q, v.data(), v.size(), cls[i]);
        }
        if (s != RK_OK) return s;
    }
    if (ds->nu == 0 || nq == 0) return RK_OK;
    rk_ctx* c = ds->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    // every query's records, one upload
    std::vector<size_t> roff(nq);
    size_t total = 0;
    for (size_t i = 0; i < nq; ++i) {
        roff[i] = total;
        total += cls[i].size();
    }
    RK_HIP(c, ds->h_rec.reserve(total));
    RK_HIP(c, ds->h_orig.reserve(total));
    RK_HIP(c, ds->rec.reserve(total));
    RK_HIP(c, ds->orig.reserve(total));
    for (size_t i = 0; i < nq; ++i) {
        const uint32_t q = qs[i].q;
        const float qx = ds->hfx[q], qy = ds->hfy[q];
        double wx, wy;
        axis_bounds(ds->ubox[0], ds->ubox[1], ds->ubox[2], ds->ubox[3], qx, qy, &wx, &wy);
        for (size_t j = 0; j < cls[i].size(); ++j) {
            const uint32_t f = cls[i][j];
            ds->h_rec.p[roff[i] + j] = make_record(qx, qy, ds->hfx[f], ds->hfy[f], wx, wy);
            ds->h_orig.p[roff[i] + j] = make_float2(ds->hfx[f], ds->hfy[f]);
        }
    }
    if (total) {
        RK_HIP(c, hipMemcpyAsync(ds->rec.p, ds->h_rec.p, total * sizeof(CandRec), hipMemcpyHostToDevice, c->stream));
        RK_HIP(c, hipMemcpyAsync(ds->orig.p, ds->h_orig.p, total * sizeof(float2), hipMemcpyHostToDevice, c->stream));
    }
    RK_HIP(c, ds->ctr.reserve(kPlainCounters));
    RK_HIP(c, ds->h_ctr.reserve(kPlainCounters));
    RK_HIP(c, hipMemsetAsync(ds->ctr.p, 0, kPlainCounters * sizeof(uint32_t), c->stream));
    RK_HIP(c, ds->counts.reserve(ds->n_pad));
    if (!sorted_out) RK_HIP(c, ds->counts_orig.reserve(ds->nu));
    std::vector<QDev> qd(nq);
    std::vector<int> var(nq), width(nq);
    for (size_t i = 0; i < nq; ++i) {
        qd[i] = make_qdev(ds, qs[i].q, 1, cls[i].size());
        width[i] = sorted_out ? 4 : count_bytes(qd[i].nc);
        rk_status s = pick_count_variant(ds, qd[i], ds->rec.p + roff[i], ds->orig.p + roff[i], ds->counts.p,
                                         sorted_out, &var[i]);
        if (s != RK_OK) return s;
    }
    // the launches: consecutive queries whose records fit the LDS staging
    // together are counted in one launch (k_count_set: each wave's users
    // loaded once for the whole set); a longer list, or every query when
    // RK_COUNT_SET=1, gets a launch of its own.  The queries of a set share
    // its widest count type and its first query's count variant (every
    // variant gives the same counts).
    size_t set_max = kCountSetMax;
    if (const char* e = std::getenv("RK_COUNT_SET")) {
        const long v = std::strtol(e, nullptr, 10);
        set_max = v < 1 ? 1 : v > (long)kCountSetMax ? kCountSetMax : (size_t)v;
    }
    struct Launch {
        size_t i0, i1;
        uint32_t n_recs;
    };
    // (a set's count slices also fit one round, kCountManyBytes)
    auto slice_bytes = [&](int w) { return ((size_t)ds->n_pad * w + 255) & ~size_t(255); };
    std::vector<Launch> launches;
    for (size_t i = 0; i < nq;) {
        Launch L{i, i + 1, qd[i].nc};
        if (qd[i].nc <= kCountSetRecs) {
            int w = width[i];
            while (L.i1 < nq && L.i1 - L.i0 < set_max && L.n_recs + qd[L.i1].nc <= kCountSetRecs &&
                   (L.i1 - L.i0 + 1) * slice_bytes(std::max(w, width[L.i1])) <= kCountManyBytes) {
                w = std::max(w, width[L.i1]);
                L.n_recs += qd[L.i1].nc;
                ++L.i1;
            }
            for (size_t j = L.i0; j < L.i1; ++j) width[j] = w;
        }
        launches.push_back(L);
        i = L.i1;
    }
    RK_HIP(c, ds->setq.reserve(nq));
    RK_HIP(c, ds->h_setq.reserve(nq));
    // counts into the context device's memory: in caller order the scatter
    // writes the caller's array; in the dataset's order the count kernels do
    const bool to_device = !sorted_out && device_dst(counts, c->device);
    const bool direct = sorted_out && device_dst(counts, c->device);
    const double t1 = now_s();
    double kernel_s = 0, out_s = 0;
    uint64_t n_launches = 0;
    for (size_t l0 = 0; l0 < launches.size();) {
        // this round: the launches whose count slices fit kCountManyBytes
        const size_t i0 = launches[l0].i0;
        std::vector<size_t> soff;
        size_t used = 0, l1 = l0;
        while (l1 < launches.size()) {
            size_t b = 0;
            for (size_t i = launches[l1].i0; i < launches[l1].i1; ++i)
                b += ((size_t)ds->n_pad * width[i] + 255) & ~size_t(255);
            if (l1 > l0 && used + b > kCountManyBytes) break;
            for (size_t i = launches[l1].i0; i < launches[l1].i1; ++i) {
                soff.push_back(used);
                used += ((size_t)ds->n_pad * width[i] + 255) & ~size_t(255);
            }
            ++l1;
        }
        const size_t i1 = launches[l1 - 1].i1;
        if (!direct) RK_HIP(c, ds->counts_many.reserve(used));
        // a query's counts: its slice of the round, or its row of the caller's array
        unsigned char* const obase = direct ? reinterpret_cast<unsigned char*>(counts) : ds->counts_many.p;
        auto out_off = [&](size_t i) -> uint64_t { // in counts of width[i]
            return direct ? (uint64_t)i * stride : soff[i - i0] / width[i];
        };
        // the sets' descriptors, one upload for the round
        for (size_t i = i0; i < i1; ++i) {
            CountSetQ& d = ds->h_setq.p[i - i0];
            d = CountSetQ{};
            d.qx = qd[i].qx;
            d.qy = qd[i].qy;
            d.nc = qd[i].nc;
            d.coff = out_off(i);
        }
        for (size_t l = l0; l < l1; ++l)
            for (size_t i = launches[l].i0; i < launches[l].i1; ++i)
                ds->h_setq.p[i - i0].roff = (uint32_t)(roff[i] - roff[launches[l].i0]);
        RK_HIP(c, hipMemcpyAsync(ds->setq.p, ds->h_setq.p, (i1 - i0) * sizeof(CountSetQ), hipMemcpyHostToDevice,
                                 c->stream));
        // the count launches back to back, one event pair around them
        RK_HIP(c, hipEventRecord(c->ev0, c->stream));
        for (size_t l = l0; l < l1; ++l) {
            const Launch& L = launches[l];
            if (L.i1 - L.i0 == 1) {
                launch_count_all(ds->view, qd[L.i0], ds->rec.p + roff[L.i0], ds->orig.p + roff[L.i0],
                                 obase + out_off(L.i0) * width[L.i0], ds->ctr.p + 1, var[L.i0], c->stream,
                                 sorted_out);
            } else {
                // every query's counts are addressed from the round's base
                launch_count_set(ds->view, ds->setq.p + (L.i0 - i0), (uint32_t)(L.i1 - L.i0), L.n_recs,
                                 ds->rec.p + roff[L.i0], ds->orig.p + roff[L.i0], obase, width[L.i0],
                                 ds->ctr.p + 1, var[L.i0], c->stream);
            }
        }
        n_launches += l1 - l0;
        RK_HIP(c, hipGetLastError());
        RK_HIP(c, hipEventRecord(c->ev1, c->stream));
        RK_HIP(c, hipEventSynchronize(c->ev1));
        float ms = 0.f;
        RK_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        kernel_s += ms * 1e-3;
        const double d0 = now_s();
        for (size_t i = i0; i < i1 && !direct; ++i) {
            const void* slice = ds->counts_many.p + soff[i - i0];
            rk_status s;
            if (sorted_out) {
                s = download(c, counts + i * stride, slice, ds->nu * sizeof(uint32_t));
            } else if (to_device) { // the scatter writes the caller's device array
                launch_scatter_counts(slice, width[i], ds->perm.p, (uint32_t)ds->nu, counts + i * stride, c->stream);
                RK_HIP(c, hipGetLastError());
                s = RK_OK;
            } else {
                launch_scatter_counts(slice, width[i], ds->perm.p, (uint32_t)ds->nu, ds->counts_orig.p, c->stream);
                RK_HIP(c, hipGetLastError());
                s = download(c, counts + i * stride, ds->counts_orig.p, ds->nu * sizeof(uint32_t));
            }
            if (s != RK_OK) return s;
        }
        if (to_device) RK_HIP(c, hipStreamSynchronize(c->stream));
        out_s += now_s() - d0;
        l0 = l1;
    }
    if (t) {
        RK_HIP(c, hipMemcpyAsync(ds->h_ctr.p, ds->ctr.p, kPlainCounters * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 c->stream));
        RK_HIP(c, hipStreamSynchronize(c->stream));
        uint64_t fp64 = 0;
        for (int j = 1; j < kPlainCounters; ++j) fp64 += ds->h_ctr.p[j];
        t->filter_s = filter_s;
        t->upload_s = t1 - t0 - filter_s;
        t->kernel_s = kernel_s;
        t->download_s = out_s;
        t->n_candidates = total;
        t->n_fp64_rechecks = fp64;
        t->n_launches = n_launches;
        t->total_s = now_s() - t0;
    }
    return RK_OK;
}

rk_status count_all_impl(rk_dataset* ds, uint32_t q, const uint32_t* cand, size_t nc, uint32_t* counts,
                         rk_timing* t, bool sorted_out)
{
    const double t0 = now_s();
    rk_status s = check_query(ds, q, 1);
    if (s != RK_OK) return s;
    if (ds->nu && !counts) return fail(ds->ctx, RK_ERR_INVALID_ARG, "NULL output");
    std::vector<uint32_t> cl;
    s = clean_candidates(ds, q, cand, nc, cl);
    if (s != RK_OK) return s;
    if (t) *t = rk_timing{};
    if (ds->nu == 0) return RK_OK;
    rk_ctx* c = ds->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RK_HIP(c, hipSetDevice(c->device));
    const double t1 = now_s();
    s = upload_records(ds, q, cl);
    if (s != RK_OK) return s;
    RK_HIP(c, ds->counts.reserve(ds->n_pad));
    if (!sorted_out) RK_HIP(c, ds->counts_orig.reserve(ds->nu));
    // [0] unused, [1 .. kBuckets] binary64 tests
    RK_HIP(c, ds->ctr.reserve(kPlainCounters));
    RK_HIP(c, ds->h_ctr.reserve(kPlainCounters));
    RK_HIP(c, hipMemsetAsync(ds->ctr.p, 0, kPlainCounters * sizeof(uint32_t), c->stream));
    QDev qd = make_qdev(ds, q, 1, cl.size());
    int variant = 0;
    s = pick_count_variant(ds, qd, ds->rec.p, ds->orig.p, ds->counts.p, sorted_out, &variant);
    if (s != RK_OK) return s;
    // the dataset's order into the context device's memory: the kernel writes
    // the caller's array
    const bool direct = sorted_out && device_dst(counts, c->device);
    RK_HIP(c, hipEventRecord(c->ev0, c->stream));
    launch_count_all(ds->view, qd, ds->rec.p, ds->orig.p, direct ? (void*)counts : (void*)ds->counts.p, ds->ctr.p + 1,
                     variant, c->stream, sorted_out);
    RK_HIP(c, hipGetLastError());
    RK_HIP(c, hipEventRecord(c->ev1, c->stream));
    if (t) RK_HIP(c, hipEventSynchronize(c->ev1)); // download_s starts when the counts are ready
    const double t2 = now_s();
    // caller order into device memory: the scatter writes the caller's array
    const bool to_device = !sorted_out && device_dst(counts, c->device);
    if (!sorted_out) {
        launch_scatter_counts(ds->counts.p, count_bytes(qd.nc), ds->perm.p, (uint32_t)ds->nu,
                              to_device ? counts : ds->counts_orig.p, c->stream);
        RK_HIP(c, hipGetLastError());
    }
    RK_HIP(c, hipMemcpyAsync(ds->h_ctr.p, ds->ctr.p, kPlainCounters * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->stream));
    if (to_device || direct) {
        RK_HIP(c, hipStreamSynchronize(c->stream));
    } else {
        s = download(c, counts, sorted_out ? ds->counts.p : ds->counts_orig.p, ds->nu * sizeof(uint32_t));
        if (s != RK_OK) return s;
    }
    float ms = 0.f;
    RK_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (t) {
        uint64_t fp64 = 0;
        for (int j = 1; j < kPlainCounters; ++j) fp64 += ds->h_ctr.p[j];
        const double t3 = now_s();
        t->kernel_s = ms * 1e-3;
        t->download_s = t3 - t2; // the scatter to caller order, if asked, and the copy of nu counts
        t->n_candidates = cl.size();
        t->n_fp64_rechecks = fp64;
        t->n_launches = 1;
        t->upload_s = 0;
        t->total_s = t3 - t0;
        t->filter_s = t1 - t0;
    }
    return RK_OK;
}

} // namespace
