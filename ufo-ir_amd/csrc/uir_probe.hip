// uir_probe.hip -- measured throughput ceilings of the two resources that
// bound the projectors: LDS reads and FP32 fused multiply-adds.
//
// The Joseph FP and BP read one 8-byte (e, d) pair from LDS per sample and
// issue about five vector operations per sample, so their roofline is set by
// the LDS read rate (bytes per clock per CU) and the vector issue rates: FP32
// adds and FMAs (wave64 instructions per clock per CU), and conversions,
// which issue at a lower rate.  None of these is a published figure for this
// build's instruction mix (no packed FP32).  These loops measure them on the
// device at hand.
#include <hip/hip_runtime.h>

#include "uir_device.h"

namespace {

constexpr int kProbeLds = 4096; // float2 entries: 32 KB of LDS per workgroup
constexpr int kProbeUnroll = 16;

// Each thread reads 16 pairs per iteration at addresses 128 entries apart
// from a per-iteration base below kProbeLds / 2, so every index is below
// kProbeLds / 2 + 15 * 128 < kProbeLds.  Lanes read consecutive entries.
// Every read is one 8-byte instruction, as in the projectors, whose addresses
// are data-dependent: m (0 at run time) keeps the compiler from merging two
// reads into one two-address instruction, which moves fewer bytes per clock.
__global__ __launch_bounds__(256) void k_probe_lds(float *__restrict__ out, int iters, float seed, int m)
{
    __shared__ float2 s[kProbeLds];
    for (int i = threadIdx.x; i < kProbeLds; i += 256)
        s[i] = make_float2(seed * (float)i, seed + (float)i);
    __syncthreads();
    const int base = (int)threadIdx.x; // < 256
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    for (int it = 0; it < iters; ++it) {
        const int b0 = (base + it * 64) & (kProbeLds / 2 - 1);
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) {
            const float2 v = s[((b0 + k * 128) ^ m) & (kProbeLds - 1)];
            if (k & 1) {
                a0 += v.x;
                a1 += v.y;
            } else {
                a2 += v.x;
                a3 += v.y;
            }
        }
    }
    out[blockIdx.x * 256 + threadIdx.x] = a0 + a1 + a2 + a3;
}

// Eight independent FMA chains per thread, 16 x 8 FMAs per iteration.
__global__ __launch_bounds__(256) void k_probe_fma(float *__restrict__ out, int iters, float seed)
{
    float a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
        a[j] = seed * (float)(threadIdx.x + j);
    const float m = seed * 0.5f, c = seed * 0.25f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                a[j] = __builtin_fmaf(a[j], m, c);
    }
    float t = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        t += a[j];
    out[blockIdx.x * 256 + threadIdx.x] = t;
}

// Eight independent chains of float-to-int conversions with floor (the
// projectors' other VALU class: conversions, fract and shift-adds issue at
// half the rate of float adds and FMAs).  Each operand is the previous
// result's bits read as a float.
__global__ __launch_bounds__(256) void k_probe_cvt(float *__restrict__ out, int iters, float seed)
{
    int b[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
        b[j] = __float_as_int(seed * (float)(threadIdx.x + j));
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                asm volatile("v_cvt_flr_i32_f32 %0, %1" : "=v"(b[j]) : "v"(b[j]));
    }
    int t = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        t += b[j];
    out[blockIdx.x * 256 + threadIdx.x] = (float)t;
}

} // namespace

int uirk_probe_ceilings(void *stream, uirk_ceilings *c)
{
    if (!c)
        return 1;
    hipStream_t st = (hipStream_t)stream;
    int dev = 0, cus = 0, khz = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, dev);
    if (e != hipSuccess || cus <= 0 || khz <= 0)
        return 3;
    const int blocks = cus * 8; // 8 workgroups of 256 threads per CU
    float *out = nullptr;
    if (hipMalloc((void **)&out, sizeof(float) * (size_t)blocks * 256) != hipSuccess)
        return 2;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = 0;
    float ms_lds = 0.0f, ms_fma = 0.0f, ms_cvt = 0.0f;
    const int lds_iters = 4096, fma_iters = 2048, reps = 5;
    if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) {
        rc = 3;
    } else {
        // warm-up launch of each, then reps timed launches
        hipLaunchKernelGGL(k_probe_lds, dim3(blocks), dim3(256), 0, st, out, lds_iters, 1.0f, 0);
        if (hipEventRecord(ev[0], st) != hipSuccess)
            rc = 3;
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_probe_lds, dim3(blocks), dim3(256), 0, st, out, lds_iters, 1.0f, 0);
        if (hipEventRecord(ev[1], st) != hipSuccess)
            rc = 3;
        if (rc == 0 &&
            (hipEventSynchronize(ev[1]) != hipSuccess || hipEventElapsedTime(&ms_lds, ev[0], ev[1]) != hipSuccess))
            rc = 3;
        hipLaunchKernelGGL(k_probe_fma, dim3(blocks), dim3(256), 0, st, out, fma_iters, 1.0f);
        if (hipEventRecord(ev[0], st) != hipSuccess)
            rc = 3;
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_probe_fma, dim3(blocks), dim3(256), 0, st, out, fma_iters, 1.0f);
        if (hipEventRecord(ev[1], st) != hipSuccess)
            rc = 3;
        if (rc == 0 &&
            (hipEventSynchronize(ev[1]) != hipSuccess || hipEventElapsedTime(&ms_fma, ev[0], ev[1]) != hipSuccess))
            rc = 3;
        hipLaunchKernelGGL(k_probe_cvt, dim3(blocks), dim3(256), 0, st, out, fma_iters, 1.0f);
        if (hipEventRecord(ev[0], st) != hipSuccess)
            rc = 3;
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_probe_cvt, dim3(blocks), dim3(256), 0, st, out, fma_iters, 1.0f);
        if (hipEventRecord(ev[1], st) != hipSuccess)
            rc = 3;
        if (rc == 0 &&
            (hipEventSynchronize(ev[1]) != hipSuccess || hipEventElapsedTime(&ms_cvt, ev[0], ev[1]) != hipSuccess))
            rc = 3;
        if (rc == 0 && hipGetLastError() != hipSuccess)
            rc = 3;
    }
    for (int i = 0; i < 2; ++i)
        if (ev[i])
            (void)hipEventDestroy(ev[i]);
    (void)hipFree(out);
    if (rc)
        return rc;
    const double clocks_lds = (double)ms_lds / reps * (double)khz; // kHz = clocks per ms
    const double clocks_fma = (double)ms_fma / reps * (double)khz;
    const double clocks_cvt = (double)ms_cvt / reps * (double)khz;
    const double lds_bytes = (double)blocks * 256 * lds_iters * kProbeUnroll * 8.0;
    const double fma_wave = (double)blocks * 4 * fma_iters * kProbeUnroll * 8.0;
    c->n_cu = cus;
    c->clock_ghz = khz / 1e6;
    c->lds_bytes_per_clock_cu = lds_bytes / cus / clocks_lds;
    c->fma_wave64_per_clock_cu = fma_wave / cus / clocks_fma;
    c->cvt_wave64_per_clock_cu = fma_wave / cus / clocks_cvt; // same instruction count
    c->lds_ms = ms_lds / reps;
    c->fma_ms = ms_fma / reps;
    return 0;
}
