// uir_asd.hip -- the per-slice steps of batched ASD-POCS (ufo-ir-asdpocs-task.c:411-509)
// between the projections: positivity, the image distances, the TV steps
// and the FP64 reductions that set the TV step length and the data-constraint
// test, all on the device so that one iteration replays as a hipGraph.
//
// Every kernel runs over a grid (blocks per slice, slices): block (c, s)
// handles elements [c * kChunk, (c + 1) * kChunk) of slice s and writes one
// FP64 partial sum; uirk_asd_finish adds a slice's partials in a fixed order
// and applies the step's scalar logic to the slice's state.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "uir_device.h"

namespace {

constexpr int kThreads = 256;
constexpr uint32_t kChunk = UIRK_ASD_CHUNK;
constexpr int kPer = kChunk / kThreads;

__device__ __forceinline__ float ld(const float *p, uint32_t i) { return p[i]; }
__device__ __forceinline__ float ld(const __half *p, uint32_t i) { return __half2float(p[i]); }
__device__ __forceinline__ float ld(const __hip_bfloat16 *p, uint32_t i) { return __bfloat162float(p[i]); }
__device__ __forceinline__ void st(float *p, uint32_t i, float v) { p[i] = v; }
__device__ __forceinline__ void st(__half *p, uint32_t i, float v) { p[i] = __float2half(v); }
__device__ __forceinline__ void st(__hip_bfloat16 *p, uint32_t i, float v) { p[i] = __float2bfloat16(v); }

// Block sum of one double per thread; thread 0 writes it to *out.
__device__ __forceinline__ void block_sum(double v, double *out)
{
    __shared__ double red[kThreads / 64];
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < kThreads / 64; ++w)
            s += red[w];
        *out = s;
    }
}

__device__ __forceinline__ double measure(float d, int l2) { return l2 ? (double)d * (double)d : fabs((double)d); }

// mode 0: x = 0 (the reference restarts its df-minimizer from zero,
// ufo-ir-sart-task.c:205-207); 1: x_prev = x (textbook: warm start, dp
// measured from the image before the data step); 2: x = x_prev (the
// reference's output, the image before the TV steps, :468); 3: x = g read as
// image-dtype data (the SART step's second buffer).
template <typename TI>
__global__ __launch_bounds__(kThreads) void k_asd_begin(const uirk_asd_slice *__restrict__ sl, int mode)
{
    const uirk_asd_slice S = sl[blockIdx.y];
    const uint32_t nv = S.n * S.n, base = blockIdx.x * kChunk;
    if (base >= nv)
        return;
    TI *x = (TI *)S.x, *xp = (TI *)S.x_prev;
    for (int k = 0; k < kPer; ++k) {
        const uint32_t i = base + k * kThreads + threadIdx.x;
        if (i < nv) {
            if (mode == 1)
                xp[i] = x[i];
            else if (mode == 2)
                x[i] = xp[i];
            else if (mode == 3)
                x[i] = ((const TI *)S.g)[i];
            else
                st(x, i, 0.0f);
        }
    }
}

// After the data step: x = POSC(x) (ufo-ir-asdpocs-task.c:463-465), partial of
// |x - x_prev| (dp, :481-483), then x_prev = x (:490).
template <typename TI>
__global__ __launch_bounds__(kThreads) void k_asd_posc_dp(const uirk_asd_slice *__restrict__ sl, int posc, int l2,
                                                          double *__restrict__ partials, int nblk)
{
    const uirk_asd_slice S = sl[blockIdx.y];
    const uint32_t nv = S.n * S.n, base = blockIdx.x * kChunk;
    double acc = 0;
    if (base < nv) {
        TI *x = (TI *)S.x, *xp = (TI *)S.x_prev;
        for (int k = 0; k < kPer; ++k) {
            const uint32_t i = base + k * kThreads + threadIdx.x;
            if (i < nv) {
                float v = ld(x, i);
                if (posc && !(v > 0.0f)) {
                    v = 0.0f;
                    st(x, i, v);
                }
                acc += measure(v - ld(xp, i), l2);
                st(xp, i, v);
            }
        }
    }
    block_sum(acc, &partials[(size_t)blockIdx.y * nblk + blockIdx.x]);
}

// Partial of |x - x_prev| (dg, :496-497).
template <typename TI>
__global__ __launch_bounds__(kThreads) void k_asd_diff(const uirk_asd_slice *__restrict__ sl, int l2,
                                                       double *__restrict__ partials, int nblk)
{
    const uirk_asd_slice S = sl[blockIdx.y];
    const uint32_t nv = S.n * S.n, base = blockIdx.x * kChunk;
    double acc = 0;
    if (base < nv) {
        const TI *x = (const TI *)S.x, *xp = (const TI *)S.x_prev;
        for (int k = 0; k < kPer; ++k) {
            const uint32_t i = base + k * kThreads + threadIdx.x;
            if (i < nv)
                acc += measure(ld(x, i) - ld(xp, i), l2);
        }
    }
    block_sum(acc, &partials[(size_t)blockIdx.y * nblk + blockIdx.x]);
}

// Partial of |r| over the residual sinogram b - A x (dd, :471-478).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_asd_sino(const uirk_asd_slice *__restrict__ sl, int l2,
                                                       double *__restrict__ partials, int nblk)
{
    const uirk_asd_slice S = sl[blockIdx.y];
    const uint32_t ns = S.n_sino;
    double acc = 0;
    for (uint32_t base = blockIdx.x * kChunk; base < ns; base += (uint32_t)nblk * kChunk) {
        const T *r = (const T *)S.r;
        for (int k = 0; k < kPer; ++k) {
            const uint32_t i = base + k * kThreads + threadIdx.x;
            if (i < ns)
                acc += measure(ld(r, i), l2);
        }
    }
    block_sum(acc, &partials[(size_t)blockIdx.y * nblk + blockIdx.x]);
}

template <typename TI> __device__ __forceinline__ float nb(const TI *img, int n, int x, int y)
{
    if (x < 0 || y < 0 || x >= n || y >= n)
        return 0.0f;
    return ld(img, (uint32_t)y * (uint32_t)n + (uint32_t)x);
}

// g = l1_grad(x) (ufo-math-tvstd-method.cl:5-57: smoothed TV gradient, eps
// 1e-3, zero outside the image) and the partial of |g| for this TV step;
// nothing once the slice's TV steps have ended.  Block c of a slice covers
// columns [256 (c % cx), +256) x rows [kTvRows (c / cx), +kTvRows): each thread
// walks down one column with the 3 x 3 neighbourhood in registers, loading
// one row (x - 1, x, x + 1) per pixel, all coalesced.  The three quotients
// v / sqrt(w) are formed as v * rsqrt(w) (hardware reciprocal square root,
// ~1 ulp).
template <typename TI>
__global__ __launch_bounds__(kThreads) void k_asd_tvgrad(const uirk_asd_slice *__restrict__ sl,
                                                         const uirk_asd_state *__restrict__ state, int l2,
                                                         double *__restrict__ partials, int nblk)
{
    const uirk_asd_slice S = sl[blockIdx.y];
    const int n = (int)S.n;
    const int cx = (n + kThreads - 1) / kThreads, cy = (n + UIRK_ASD_TV_ROWS - 1) / UIRK_ASD_TV_ROWS;
    double acc = 0;
    if ((int)blockIdx.x < cx * cy && !state[blockIdx.y].stop) {
        const TI *img = (const TI *)S.x;
        const int x = (int)(blockIdx.x % cx) * kThreads + threadIdx.x;
        const int y0 = (int)(blockIdx.x / cx) * UIRK_ASD_TV_ROWS;
        const int y1 = min(y0 + UIRK_ASD_TV_ROWS, n);
        if (x < n) {
            const float eps = 0.001f;
            // rows y - 1 (a), y (b); columns x - 1 (m), x (0), x + 1 (p)
            float am = nb(img, n, x - 1, y0 - 1), a0 = nb(img, n, x, y0 - 1);
            float bm = nb(img, n, x - 1, y0), b0 = nb(img, n, x, y0), bp = nb(img, n, x + 1, y0);
            for (int y = y0; y < y1; ++y) {
                const float cm = nb(img, n, x - 1, y + 1), c0 = nb(img, n, x, y + 1), cp = nb(img, n, x + 1, y + 1);
                const float f0 = am, f1 = a0, f2 = bm, f3 = b0, f4 = bp, f5 = c0, f6 = cp;
                const float w0 = eps + (f3 - f2) * (f3 - f2) + (f3 - f5) * (f3 - f5);
                const float w1 = eps + (f4 - f3) * (f4 - f3) + (f4 - f6) * (f4 - f6);
                const float w2 = eps + (f1 - f3) * (f1 - f3) + (f1 - f0) * (f1 - f0);
                const float df = (2.0f * f3 - f2 - f5) * __builtin_amdgcn_rsqf(w0) -
                                 (f4 - f3) * __builtin_amdgcn_rsqf(w1) - (f1 - f3) * __builtin_amdgcn_rsqf(w2);
                S.g[(uint32_t)y * (uint32_t)n + (uint32_t)x] = df;
                acc += measure(df, l2);
                am = bm;
                a0 = b0;
                bm = cm;
                b0 = c0;
                bp = cp;
            }
        }
    }
    block_sum(acc, &partials[(size_t)blockIdx.y * nblk + blockIdx.x]);
}

// x = x - factor * g (operation_deduction2, ufo-ir-basic-ops.cl:104-125),
// factor = dtgv / |g| from the slice's state.
template <typename TI>
__global__ __launch_bounds__(kThreads) void k_asd_tvstep(const uirk_asd_slice *__restrict__ sl,
                                                         const uirk_asd_state *__restrict__ state)
{
    const uirk_asd_slice S = sl[blockIdx.y];
    const uint32_t nv = S.n * S.n, base = blockIdx.x * kChunk;
    if (base >= nv || state[blockIdx.y].stop)
        return;
    const float f = state[blockIdx.y].factor;
    TI *x = (TI *)S.x;
    for (int k = 0; k < kPer; ++k) {
        const uint32_t i = base + k * kThreads + threadIdx.x;
        if (i < nv)
            st(x, i, ld(x, i) - f * S.g[i]);
    }
}

// One block per slice: the slice's partials added in a fixed order, then the
// step's scalar update (ufo-ir-asdpocs-task.c:478-503, tvstd :573-579).
__global__ __launch_bounds__(kThreads) void k_asd_finish(uirk_asd_state *__restrict__ state,
                                                         const double *__restrict__ partials, int nblk, int what,
                                                         uirk_asd_params p, float *__restrict__ diag, int n_slices,
                                                         int max_iters)
{
    const int s = blockIdx.x;
    double v = 0;
    for (int c = threadIdx.x; c < nblk; c += kThreads)
        v += partials[(size_t)s * nblk + c];
    __shared__ double total;
    block_sum(v, &total);
    __syncthreads();
    if (threadIdx.x != 0)
        return;
    const double norm = p.textbook ? sqrt(total) : total;
    uirk_asd_state &S = state[s];
    switch (what) {
    case UIRK_ASD_DD:
        S.dd = (float)norm;
        break;
    case UIRK_ASD_DP:
        S.dp = (float)norm;
        if (S.it == 0) // :485-487
            S.dtgv = p.alpha * S.dp;
        S.stop = 0;
        break;
    case UIRK_ASD_G:
        if (S.stop)
            break;
        if (p.textbook && !(norm > 0)) { // a zero gradient ends the steps
            S.stop = 1;
            break;
        }
        S.factor = S.dtgv / (float)norm;
        break;
    default: { // UIRK_ASD_DG: end of the iteration
        S.dg = (float)norm;
        if (S.it < max_iters) {
            float *d = diag + ((size_t)S.it * n_slices + s) * 4;
            d[0] = S.dd;
            d[1] = S.dp;
            d[2] = S.dg;
            d[3] = S.dtgv;
        }
        S.beta *= p.beta_red;                         // :498
        if (S.dg > p.r_max * S.dp && S.dd > 0.001f)  // :495, 501-503
            S.dtgv *= p.alpha_red;
        if (S.beta >= 0.0f && S.beta <= 1.0f)        // :459, a [0, 1] property
            S.relax = S.beta;
        S.it += 1;
        S.stop = 0;
    }
    }
}

template <typename TI>
int launch_image(int kind, const uirk_asd_slice *sl, int n_slices, int nblk, const uirk_asd_state *state, int flag,
                 int l2, double *partials, hipStream_t stm)
{
    const dim3 grid(nblk, n_slices);
    switch (kind) {
    case 0: hipLaunchKernelGGL(k_asd_begin<TI>, grid, dim3(kThreads), 0, stm, sl, flag); break;
    case 1: hipLaunchKernelGGL(k_asd_posc_dp<TI>, grid, dim3(kThreads), 0, stm, sl, flag, l2, partials, nblk); break;
    case 2: hipLaunchKernelGGL(k_asd_diff<TI>, grid, dim3(kThreads), 0, stm, sl, l2, partials, nblk); break;
    case 3: hipLaunchKernelGGL(k_asd_tvgrad<TI>, grid, dim3(kThreads), 0, stm, sl, state, l2, partials, nblk); break;
    default: hipLaunchKernelGGL(k_asd_tvstep<TI>, grid, dim3(kThreads), 0, stm, sl, state); break;
    }
    return (int)hipGetLastError();
}

} // namespace

extern "C" {

int uirk_asd_image(int kind, int img_dtype, const uirk_asd_slice *d_sl, int n_slices, int nblk,
                   const uirk_asd_state *d_state, int flag, int l2, double *d_partials, void *stream)
{
    if (n_slices <= 0 || nblk <= 0 || n_slices > 65535)
        return n_slices == 0 ? 0 : (int)hipErrorInvalidValue;
    hipStream_t stm = (hipStream_t)stream;
    switch (img_dtype) {
    case UIR_DTYPE_F32: return launch_image<float>(kind, d_sl, n_slices, nblk, d_state, flag, l2, d_partials, stm);
    case UIR_DTYPE_F16: return launch_image<__half>(kind, d_sl, n_slices, nblk, d_state, flag, l2, d_partials, stm);
    case UIR_DTYPE_BF16:
        return launch_image<__hip_bfloat16>(kind, d_sl, n_slices, nblk, d_state, flag, l2, d_partials, stm);
    default: return (int)hipErrorInvalidValue;
    }
}

int uirk_asd_sino_norm(int dtype, const uirk_asd_slice *d_sl, int n_slices, int nblk, int l2, double *d_partials,
                       void *stream)
{
    if (n_slices <= 0 || nblk <= 0 || n_slices > 65535)
        return n_slices == 0 ? 0 : (int)hipErrorInvalidValue;
    const dim3 grid(nblk, n_slices);
    hipStream_t stm = (hipStream_t)stream;
    switch (dtype) {
    case UIR_DTYPE_F32:
        hipLaunchKernelGGL(k_asd_sino<float>, grid, dim3(kThreads), 0, stm, d_sl, l2, d_partials, nblk);
        break;
    case UIR_DTYPE_F16:
        hipLaunchKernelGGL(k_asd_sino<__half>, grid, dim3(kThreads), 0, stm, d_sl, l2, d_partials, nblk);
        break;
    case UIR_DTYPE_BF16:
        hipLaunchKernelGGL(k_asd_sino<__hip_bfloat16>, grid, dim3(kThreads), 0, stm, d_sl, l2, d_partials, nblk);
        break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

int uirk_asd_finish(uirk_asd_state *d_state, const double *d_partials, int n_slices, int nblk, int what,
                    const uirk_asd_params *p, float *d_diag, int max_iters, void *stream)
{
    if (n_slices <= 0)
        return 0;
    hipLaunchKernelGGL(k_asd_finish, dim3(n_slices), dim3(kThreads), 0, (hipStream_t)stream, d_state, d_partials, nblk,
                       what, *p, d_diag, n_slices, max_iters);
    return (int)hipGetLastError();
}

} // extern "C"
