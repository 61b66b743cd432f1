// count-rates -- gfx950 instruction and loop rates behind the count-all
// kernel's design (rk_kernels.hip, count-all section).
//
// Every probe is a register or LDS loop with no global memory traffic in the
// timed part (one store per thread at the end), launched as 256 x W blocks of
// 256 threads (W waves per SIMD on 256 CUs) and timed with HIP events over one
// launch after a warm-up launch.  Rates are per SIMD (1,024 SIMDs).
//
//   count-rates            all tables
//
// Tables:
//   1. single VALU instructions: wave instructions per ns per SIMD, 8 waves
//   2. the clean test (FMA, FMA with clamp, add) with the candidate record
//      from registers, from an LDS broadcast, or from scalar loads: TFLOP/s
//      at 4 flops per test, U users per lane
//   3. v_mfma_f32_32x32x2_f32 alone (4 per iteration), FP32 FMAs alone (128)
//      and both in one loop: ns per wave-iteration per SIMD at 1-8 waves
//      (co-execution check: overlap makes the third row the larger of the
//      first two, none makes it their sum)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CK(x)                                                                                                 \
    do {                                                                                                      \
        hipError_t e_ = (x);                                                                                  \
        if (e_ != hipSuccess) {                                                                               \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);                      \
            return 1;                                                                                         \
        }                                                                                                     \
    } while (0)

namespace {

constexpr int kChains = 8; // independent dependency chains per lane
constexpr int kRounds = 8; // unrolled rounds per loop iteration

// 1. one instruction kind, kChains chains, kRounds rounds per iteration
#define RATE_KERNEL(NAME, ASM, CLOBBER)                                                                       \
    __global__ __launch_bounds__(256) void NAME(float* out, int iters)                                        \
    {                                                                                                         \
        float x[kChains];                                                                                     \
        const float y = 0.999f + threadIdx.x * 1e-9f, z = 1e-3f;                                              \
        for (int j = 0; j < kChains; ++j) x[j] = threadIdx.x + j;                                             \
        for (int i = 0; i < iters; ++i) {                                                                     \
            _Pragma("unroll") for (int r = 0; r < kRounds; ++r)                                               \
                _Pragma("unroll") for (int j = 0; j < kChains; ++j) asm volatile(ASM : "+v"(x[j]) : "v"(y), "v"(z) CLOBBER); \
        }                                                                                                     \
        float s = 0.f;                                                                                        \
        for (int j = 0; j < kChains; ++j) s += x[j];                                                          \
        out[blockIdx.x * blockDim.x + threadIdx.x] = s;                                                       \
    }
#define NO_CLOBBER
#define VCC_CLOBBER : "vcc"
RATE_KERNEL(r_fma, "v_fma_f32 %0, %1, %2, %0", NO_CLOBBER)
RATE_KERNEL(r_fma_clamp, "v_fma_f32 %0, %1, %2, %0 clamp", NO_CLOBBER)
RATE_KERNEL(r_fmac, "v_fmac_f32 %0, %1, %2", NO_CLOBBER)
RATE_KERNEL(r_mul, "v_mul_f32 %0, %1, %0", NO_CLOBBER)
RATE_KERNEL(r_add, "v_add_f32 %0, %1, %0", NO_CLOBBER)
RATE_KERNEL(r_max, "v_max_f32 %0, %1, %0", NO_CLOBBER)
RATE_KERNEL(r_med3, "v_med3_f32 %0, %0, %1, %2", NO_CLOBBER)
RATE_KERNEL(r_cmp, "v_cmp_lt_f32 vcc, %1, %0", VCC_CLOBBER)
RATE_KERNEL(r_alignbit, "v_alignbit_b32 %0, %0, %1, 31", NO_CLOBBER)

// 2. the clean test: acc += clamp(fma(wx, gx, fma(wy, gy, -sh)))
__device__ __forceinline__ float clean_test(float acc, float wx, float wy, float gx, float gy, float nsh)
{
    return acc + __builtin_amdgcn_fmed3f(fmaf(wx, gx, fmaf(wy, gy, nsh)), 0.0f, 1.0f);
}

template <int U>
__global__ __launch_bounds__(256) void t_registers(float* out, int iters)
{
    float wx[U], wy[U], acc[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        wx[j] = threadIdx.x * 0.01f + j;
        wy[j] = j - threadIdx.x * 0.02f;
        acc[j] = 0.f;
    }
    float gx = 0.3f + threadIdx.x * 1e-6f, gy = -0.2f, nsh = -0.1f;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int r = 0; r < 64; ++r) {
            // a new candidate every round, at no cost: the compiler must
            // assume the record changed, so nothing is hoisted or shared
            asm volatile("" : "+v"(gx), "+v"(gy), "+v"(nsh));
#pragma unroll
            for (int j = 0; j < U; ++j) acc[j] = clean_test(acc[j], wx[j], wy[j], gx, gy, nsh);
        }
    }
    float s = 0.f;
    for (int j = 0; j < U; ++j) s += acc[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// records broadcast from LDS: one 12-byte read (all lanes, one address) per
// candidate, as in the kernel's clean loop
template <int U>
__global__ __launch_bounds__(256) void t_lds(float* out, int iters)
{
    __shared__ float4 rec[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    rec[wave][lane] = make_float4(0.3f + lane * 1e-6f, -0.2f + lane * 1e-6f, -0.1f - lane * 1e-3f, 0.f);
    __syncthreads();
    float wx[U], wy[U], acc[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        wx[j] = threadIdx.x * 0.01f + j;
        wy[j] = j - threadIdx.x * 0.02f;
        acc[j] = 0.f;
    }
    for (int i = 0; i < iters; ++i) {
        asm volatile("" ::: "memory"); // the records are read again every pass
#pragma unroll 4
        for (int c = 0; c < 64; ++c) {
            const float4 r = rec[wave][c];
#pragma unroll
            for (int j = 0; j < U; ++j) acc[j] = clean_test(acc[j], wx[j], wy[j], r.x, r.y, r.z);
        }
    }
    float s = 0.f;
    for (int j = 0; j < U; ++j) s += acc[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// records through scalar loads (uniform address, read-only): an FMA takes one
// scalar operand, so -sh costs a move to a vector register per candidate
template <int U>
__global__ __launch_bounds__(256) void t_scalar(const float4* __restrict__ g, float* out, int iters)
{
    float wx[U], wy[U], acc[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        wx[j] = threadIdx.x * 0.01f + j;
        wy[j] = j - threadIdx.x * 0.02f;
        acc[j] = 0.f;
    }
    for (int i = 0; i < iters; ++i) {
        asm volatile("" ::: "memory");
#pragma unroll 4
        for (int c = 0; c < 64; ++c) {
            const float4 r = g[c];
#pragma unroll
            for (int j = 0; j < U; ++j) acc[j] = clean_test(acc[j], wx[j], wy[j], r.x, r.y, r.z);
        }
    }
    float s = 0.f;
    for (int j = 0; j < U; ++j) s += acc[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 3. matrix instructions and FP32 FMAs, alone and interleaved
typedef float f16v __attribute__((ext_vector_type(16)));

// kMfma independent accumulating chains (4 keep the matrix unit busy from
// one wave: each instruction waits only for its own chain's previous result)
// and kFma FP32 FMAs in kChains chains per iteration
template <int kMfma, int kFma>
__global__ __launch_bounds__(256) void t_mix(float* out, int iters)
{
    float a = threadIdx.x * 1e-3f, b = threadIdx.x * 2e-3f;
    f16v c[kMfma > 0 ? kMfma : 1];
    for (int m = 0; m < (kMfma > 0 ? kMfma : 1); ++m)
        for (int r = 0; r < 16; ++r) c[m][r] = r * 0.5f + threadIdx.x + m;
    float x[kChains];
    for (int j = 0; j < kChains; ++j) x[j] = threadIdx.x + j;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int m = 0; m < kMfma; ++m) c[m] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c[m], 0, 0, 0);
#pragma unroll
        for (int r = 0; r < kFma / kChains; ++r)
#pragma unroll
            for (int j = 0; j < kChains; ++j) x[j] = fmaf(x[j], 0.999f, 1e-3f);
    }
    float s = 0.f;
    for (int m = 0; m < kMfma; ++m) s += c[m][0] + c[m][15];
    for (int j = 0; j < kChains; ++j) s += x[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

hipEvent_t e0, e1;

// ms of one launch of f (after a warm-up launch)
template <typename F>
int time_ms(F f, float* ms)
{
    f();
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0));
    f();
    CK(hipGetLastError());
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(ms, e0, e1));
    return 0;
}

} // namespace

int main()
{
    constexpr int kCUs = 256, kSimds = 1024;
    float* out = nullptr;
    float4* g = nullptr;
    CK(hipMalloc(&out, (size_t)kCUs * 8 * 256 * sizeof(float)));
    CK(hipMalloc(&g, 64 * sizeof(float4)));
    float4 h[64];
    for (int i = 0; i < 64; ++i) h[i] = make_float4(0.3f + i * 1e-6f, -0.2f + i * 1e-6f, -0.1f - i * 1e-3f, 0.f);
    CK(hipMemcpy(g, h, sizeof h, hipMemcpyHostToDevice));
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    float ms = 0.f;

    std::printf("1. single instructions, 8 waves per SIMD (wave instructions per ns per SIMD)\n");
    {
        const int blocks = kCUs * 8, iters = 2000;
        const double winst = (double)blocks * 4 * iters * kRounds * kChains / kSimds;
        struct Row {
            const char* name;
            void (*k)(float*, int);
        } rows[] = {{"v_fma_f32", r_fma},   {"v_fma_f32 clamp", r_fma_clamp}, {"v_fmac_f32", r_fmac},
                    {"v_mul_f32", r_mul},   {"v_add_f32", r_add},             {"v_max_f32", r_max},
                    {"v_med3_f32", r_med3}, {"v_cmp_lt_f32", r_cmp},          {"v_alignbit_b32", r_alignbit}};
        for (const Row& r : rows) {
            if (time_ms([&] { r.k<<<blocks, 256>>>(out, iters); }, &ms)) return 1;
            std::printf("   %-18s %.3f\n", r.name, winst / (ms * 1e6));
        }
    }

    std::printf("2. clean test, 8 waves per SIMD (TFLOP/s at 4 flops per test)\n");
    {
        const int blocks = kCUs * 8, iters = 400;
        auto tflops = [&](int u, float t) { return 4.0 * blocks * 256.0 * iters * 64 * u / (t * 1e-3) / 1e12; };
#define T2(LABEL, U, LAUNCH)                                                                                  \
    do {                                                                                                      \
        if (time_ms([&] { LAUNCH; }, &ms)) return 1;                                                          \
        std::printf("   %-34s U=%d  %.1f\n", LABEL, U, tflops(U, ms));                                        \
    } while (0)
        T2("record in registers", 4, (t_registers<4><<<blocks, 256>>>(out, iters)));
        T2("record in registers", 8, (t_registers<8><<<blocks, 256>>>(out, iters)));
        T2("record broadcast from LDS", 2, (t_lds<2><<<blocks, 256>>>(out, iters)));
        T2("record broadcast from LDS", 4, (t_lds<4><<<blocks, 256>>>(out, iters)));
        T2("record broadcast from LDS", 6, (t_lds<6><<<blocks, 256>>>(out, iters)));
        T2("record broadcast from LDS", 8, (t_lds<8><<<blocks, 256>>>(out, iters)));
        T2("record through scalar loads", 4, (t_scalar<4><<<blocks, 256>>>(g, out, iters)));
        T2("record through scalar loads", 8, (t_scalar<8><<<blocks, 256>>>(g, out, iters)));
#undef T2
    }

    std::printf("3. v_mfma_f32_32x32x2_f32 and FP32 FMAs (ns per wave-iteration per SIMD; waves 1 2 4 8)\n");
    {
        const int iters = 4000;
        auto row = [&](const char* label, void (*k)(float*, int)) -> int {
            std::printf("   %-32s", label);
            for (int w : {1, 2, 4, 8}) {
                const int blocks = kCUs * w;
                if (time_ms([&] { k<<<blocks, 256>>>(out, iters); }, &ms)) return 1;
                const double per_simd = (double)blocks * 4 * iters / kSimds; // wave-iterations per SIMD
                std::printf(" %7.2f", ms * 1e6 / per_simd);
            }
            std::printf("\n");
            return 0;
        };
        if (row("4 mfma", t_mix<4, 0>)) return 1;
        if (row("128 fma", t_mix<0, 128>)) return 1;
        if (row("4 mfma + 128 fma", t_mix<4, 128>)) return 1;
    }
    CK(hipFree(out));
    CK(hipFree(g));
    return 0;
}
