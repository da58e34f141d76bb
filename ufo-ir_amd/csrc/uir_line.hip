// uir_line.hip -- the line projection model on gfx950: forward projection and
// backprojection with exact ray-pixel intersection lengths (the projector's
// `model` property, ufo-ir-parallel-projector-task.c:123-128,370, which
// selects the projector-parallel-<model>.cl kernels; the reference ships only
// joseph).  The model and the per-angle constants are described with
// uirk_line_angle in uir_device.h; the BP is the transpose of the FP.
//
// These kernels serve launches over one angle or a few (SART's sweep,
// projector subsets) and the line model's weights.  Projections over all
// angles run the table-driven kernels of uir_fast.hip in their line mode.
//
// FP: one ray per R lanes (R = 64: a thread per ray; R = 4: a ray over 16
// lanes, interleaved along it), so that a wave's loads coalesce: rays of
// family 0 advance one image row per sample and neighbouring rays read
// neighbouring pixels of that row (R = 64); rays of family 1 advance one
// column per sample and the lanes of one ray read neighbouring pixels of the
// row it crosses (R = 4).  A sample is two taps, x[c0] and x[c0 + 1] across
// the ray, weighted by the chord lengths at fractional offset f.
//
// BP: four neighbouring pixels per thread, 16 x 16 threads per workgroup, one
// angle after another; per angle two sinogram taps per pixel.
//
// The batched ASD-POCS of the line model runs its SART data step with the
// same per-angle arithmetic over every slice of a batch (k_sart_line_fp,
// k_sart_line_bp: one launch each per angle, the slice in blockIdx.y), on
// the batch's storage types.
//
// Built with -ffp-contract=off: fused multiply-adds are written out.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "uir_device.h"

namespace {

// Loads and stores in a storage type, as FP32.
__device__ __forceinline__ float ld(const float *p, size_t i) { return p[i]; }
__device__ __forceinline__ float ld(const __half *p, size_t i) { return __half2float(p[i]); }
__device__ __forceinline__ float ld(const __hip_bfloat16 *p, size_t i) { return __bfloat162float(p[i]); }
__device__ __forceinline__ void st(float *p, size_t i, float v) { p[i] = v; }
__device__ __forceinline__ void st(__half *p, size_t i, float v) { p[i] = __float2half(v); }
__device__ __forceinline__ void st(__hip_bfloat16 *p, size_t i, float v) { p[i] = __float2bfloat16(v); }

struct LineFpArgs {
    const float *vol;
    const float *r_sino;
    float *w_sino;
    const float *ray_w;
    const uirk_line_angle *tab;
    int n, n_dets, a0;
    float corr;
};

struct LineBpArgs {
    const float *r_vol;
    float *w_vol;
    const float *sino;
    const float *pix_w;
    const uirk_line_angle *tab;
    int n, n_dets, a0, na, posc;
    float relax;
};

// Sum over this lane's samples i = q + S k of one ray, k in [k0, k1):
// w0 x[c0] + w1 x[c0 + 1] with c0 = floor(v), f = v - c0, v = v0 + i step
// (family 0: pixel (i, c) is x[i][c]; family 1: x[c][i]).  Taps outside the
// image read a clamped address and count 0.
template <int S, int FAM, bool ALIGNED, typename TI>
__device__ __forceinline__ float fp_line_walk(const TI *__restrict__ vol, int N, float v0, float step, float k,
                                              float b0, float b1, bool tie0, int q, int k0, int k1)
{
    float acc = 0.0f;
    for (int kk = k0; kk < k1; kk += 4) {
        float t0[4], t1[4], f[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = q + S * (kk + u);
            const bool in = kk + u < k1;
            const float v = __builtin_fmaf((float)i, step, v0);
            const float fl = floorf(v);
            f[u] = v - fl;
            const int c0 = (int)fl;
            const bool ok0 = in && (unsigned)c0 < (unsigned)N;
            const bool ok1 = in && (unsigned)(c0 + 1) < (unsigned)N;
            const int ca = min(max(c0, 0), N - 1), cb = min(max(c0 + 1, 0), N - 1);
            const int ii = min(i, N - 1);
            const size_t oa = FAM == 0 ? (size_t)ii * N + ca : (size_t)ca * N + ii;
            const size_t ob = FAM == 0 ? (size_t)ii * N + cb : (size_t)cb * N + ii;
            const float xa = ld(vol, oa), xb = ld(vol, ob);
            t0[u] = ok0 ? xa : 0.0f;
            t1[u] = ok1 ? xb : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float w0, w1;
            if (ALIGNED) {
                w0 = (f[u] < 0.5f || (tie0 && f[u] == 0.5f)) ? 1.0f : 0.0f;
                w1 = 1.0f - w0;
            } else {
                w0 = __builtin_amdgcn_fmed3f(__builtin_fmaf(f[u], -k, b0), 0.0f, 1.0f);
                w1 = __builtin_amdgcn_fmed3f(__builtin_fmaf(f[u], k, b1), 0.0f, 1.0f);
            }
            acc = __builtin_fmaf(w0, t0[u], acc);
            acc = __builtin_fmaf(w1, t1[u], acc);
        }
    }
    return acc;
}

// The sum of ray s of angle L over this lane's samples (lane q of the ray's
// 64 / R lanes; the caller adds the lanes up).
template <int R, int FAM, typename TI>
__device__ __forceinline__ float fp_line_ray(const TI *__restrict__ vol, const uirk_line_angle &L, int N, int s, int q)
{
    constexpr int S = 64 / R; // lanes per ray
    const float v0 = (float)(L.fp_base + (double)s * L.fp_tk);
    const float step = L.fp_step;
    // samples whose taps can touch the image (v in [-1, N)), with a margin of
    // two samples on each side; nearly axis-parallel rays walk every sample
    int i_lo = 0, i_hi = N;
    if (step == 0.0f) {
        if (!(v0 >= -1.0f && v0 < (float)N))
            i_hi = 0;
    } else if (fabsf(step) >= 1.0f / 64.0f) {
        const double a = (-1.0 - (double)v0) / (double)step, b = ((double)N - (double)v0) / (double)step;
        const double lo = floor(fmin(a, b)) - 2.0, hi = ceil(fmax(a, b)) + 3.0;
        i_hi = (int)fmax(0.0, fmin((double)N, hi));
        i_lo = min((int)fmax(0.0, fmin((double)N, lo)), i_hi);
    }
    const int k0 = (i_lo - q + S - 1) / S, k1 = (i_hi - q + S - 1) / S;
    if (L.aligned)
        return fp_line_walk<S, FAM, true>(vol, N, v0, step, L.fp_k, L.fp_b0, L.fp_b1, L.tie0 != 0, q, k0, k1);
    return fp_line_walk<S, FAM, false>(vol, N, v0, step, L.fp_k, L.fp_b0, L.fp_b1, false, q, k0, k1);
}

template <int R, int FAM>
__global__ __launch_bounds__(256) void k_fp_line(LineFpArgs p)
{
    const int lane = threadIdx.x & 63;
    const int r = lane % R, q = lane / R;
    const int s = (blockIdx.x * 4 + (threadIdx.x >> 6)) * R + r;
    const int ang = p.a0 + blockIdx.y;
    const uirk_line_angle &L = p.tab[ang];
    const int N = p.n, D = p.n_dets;
    float acc = s < D ? fp_line_ray<R, FAM>(p.vol, L, N, s, q) : 0.0f;
#pragma unroll
    for (int m = R; m < 64; m <<= 1)
        acc += __shfl_xor(acc, m, 64);
    if (s < D && q == 0) {
        const size_t o = (size_t)ang * D + s;
        float w = p.r_sino[o] + (acc * L.hm) * p.corr;
        if (p.ray_w)
            w = w * p.ray_w[o];
        p.w_sino[o] = w;
    }
}

// SART data step of the batched ASD-POCS, FP half: for angle ang of slice
// blockIdx.y, r[ang] = (b[ang] + corr FP) * ray_w, four rays per wave of
// either family.  Slices with fewer angles skip it.
template <typename T, typename TI>
__global__ __launch_bounds__(256) void k_sart_line_fp(const uirk_sart_slice *__restrict__ slices, int ang)
{
    const uirk_sart_slice &Sl = slices[blockIdx.y];
    const int N = Sl.n, D = Sl.n_dets;
    if (ang >= Sl.n_angles)
        return; // uniform over the block
    constexpr int R = 4;
    const int lane = threadIdx.x & 63;
    const int r = lane % R, q = lane / R;
    const int s = (blockIdx.x * 4 + (threadIdx.x >> 6)) * R + r;
    const uirk_line_angle &L = Sl.line[ang];
    const TI *vol = (const TI *)Sl.x_in;
    float acc = 0.0f;
    if (s < D)
        acc = L.family == 0 ? fp_line_ray<R, 0>(vol, L, N, s, q) : fp_line_ray<R, 1>(vol, L, N, s, q);
#pragma unroll
    for (int m = R; m < 64; m <<= 1)
        acc += __shfl_xor(acc, m, 64);
    if (s < D && q == 0) {
        const size_t o = (size_t)ang * D + s;
        float w = ld((const T *)Sl.b, o) + (acc * L.hm) * Sl.corr;
        if (Sl.ray_w)
            w = w * Sl.ray_w[o];
        st((T *)Sl.r, o, w);
    }
}

__global__ __launch_bounds__(256) void k_bp_line(LineBpArgs p)
{
    const int N = p.n, D = p.n_dets;
    const int x0 = (blockIdx.x * 16 + (threadIdx.x & 15)) * 4;
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x0 >= N || y >= N)
        return; // no barriers below
    const float Y = (float)y + 0.5f - 0.5f * (float)N;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int a = p.a0; a < p.a0 + p.na; ++a) {
        const uirk_line_angle &L = p.tab[a];
        const float *row = p.sino + (size_t)a * D;
        const float rowoff = __builtin_fmaf(-Y, L.s, L.bp_off);
        float f[4], t0[4], t1[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float u = __builtin_fmaf((float)(x0 + k), L.c, rowoff);
            const float fl = floorf(u);
            f[k] = u - fl;
            const int j = (int)fl;
            const float ta = row[min(max(j, 0), D - 1)], tb = row[min(max(j + 1, 0), D - 1)];
            t0[k] = (unsigned)j < (unsigned)D ? ta : 0.0f;
            t1[k] = (unsigned)(j + 1) < (unsigned)D ? tb : 0.0f;
        }
        if (L.aligned) { // the bin whose ray offset d = -f or 1 - f is in [-1/2, 1/2)
#pragma unroll
            for (int k = 0; k < 4; ++k)
                acc[k] += f[k] <= 0.5f ? t0[k] : t1[k];
        } else {
            const float bk = L.bp_k, b0 = L.bp_b0, b1 = L.bp_b1, cap = L.hm;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float w0 = __builtin_amdgcn_fmed3f(__builtin_fmaf(f[k], -bk, b0), 0.0f, cap);
                const float w1 = __builtin_amdgcn_fmed3f(__builtin_fmaf(f[k], bk, b1), 0.0f, cap);
                acc[k] = __builtin_fmaf(w0, t0[k], acc[k]);
                acc[k] = __builtin_fmaf(w1, t1[k], acc[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int px = x0 + k;
        if (px >= N)
            break;
        const size_t o = (size_t)y * N + px;
        const float bp = p.relax * acc[k];
        float out = p.pix_w ? bp * p.pix_w[o] + p.r_vol[o] : p.r_vol[o] + bp;
        if (p.posc)
            out = out > 0.0f ? out : 0.0f;
        p.w_vol[o] = out;
    }
}

// SART data step, BP half: x += relax A_ang^T r[ang] in place for slice
// blockIdx.y (relax from the slice's state; x_in = x_out), four neighbouring
// pixels per thread and 16 x 16 threads per workgroup, as k_bp_line with one
// angle.  blockIdx.x enumerates the 64 x 16 pixel blocks of the largest
// image (tiles_x across); slices with fewer angles skip it.
template <typename T, typename TI>
__global__ __launch_bounds__(256) void k_sart_line_bp(const uirk_sart_slice *__restrict__ slices,
                                                      const uirk_asd_state *__restrict__ state, int ang, int tiles_x)
{
    const uirk_sart_slice &Sl = slices[blockIdx.y];
    const int N = Sl.n, D = Sl.n_dets;
    if (ang >= Sl.n_angles)
        return; // uniform over the block
    const int bx = (int)blockIdx.x % tiles_x, by = (int)blockIdx.x / tiles_x;
    const int x0 = (bx * 16 + (threadIdx.x & 15)) * 4;
    const int y = by * 16 + (threadIdx.x >> 4);
    if (x0 >= N || y >= N)
        return; // no barriers below
    const uirk_line_angle &L = Sl.line[ang];
    const T *row = (const T *)Sl.r + (size_t)ang * D;
    const float relax = state[blockIdx.y].relax;
    const float Y = (float)y + 0.5f - 0.5f * (float)N;
    const float rowoff = __builtin_fmaf(-Y, L.s, L.bp_off);
    const TI *xin = (const TI *)Sl.x_in;
    TI *xout = (TI *)Sl.x_out;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int px = x0 + k;
        if (px >= N)
            break;
        const float u = __builtin_fmaf((float)px, L.c, rowoff);
        const float fl = floorf(u);
        const float f = u - fl;
        const int j = (int)fl;
        const float ta = ld(row, (size_t)min(max(j, 0), D - 1)), tb = ld(row, (size_t)min(max(j + 1, 0), D - 1));
        const float t0 = (unsigned)j < (unsigned)D ? ta : 0.0f;
        const float t1 = (unsigned)(j + 1) < (unsigned)D ? tb : 0.0f;
        float acc;
        if (L.aligned) { // the bin whose ray offset d = -f or 1 - f is in [-1/2, 1/2)
            acc = f <= 0.5f ? t0 : t1;
        } else {
            const float w0 = __builtin_amdgcn_fmed3f(__builtin_fmaf(f, -L.bp_k, L.bp_b0), 0.0f, L.hm);
            const float w1 = __builtin_amdgcn_fmed3f(__builtin_fmaf(f, L.bp_k, L.bp_b1), 0.0f, L.hm);
            acc = __builtin_fmaf(w1, t1, __builtin_fmaf(w0, t0, 0.0f));
        }
        const size_t o = (size_t)y * N + px;
        st(xout, o, ld(xin, o) + relax * acc);
    }
}

template <typename T, typename TI>
int launch_sart_line(const uirk_sart_slice *sl, const uirk_asd_state *state, int n_slices, int max_n, int max_d,
                     int ang, hipStream_t stream)
{
    hipLaunchKernelGGL((k_sart_line_fp<T, TI>), dim3((max_d + 15) / 16, n_slices), dim3(256), 0, stream, sl, ang);
    const int tx = (max_n + 63) / 64, ty = (max_n + 15) / 16;
    hipLaunchKernelGGL((k_sart_line_bp<T, TI>), dim3(tx * ty, n_slices), dim3(256), 0, stream, sl, state, ang, tx);
    return (int)hipGetLastError();
}

} // namespace

extern "C" {

int uirk_sart_line(const uirk_sart_slice *d_slices, const uirk_asd_state *d_state, int n_slices, int max_n, int max_d,
                   int dtype, int img_dtype, int angle, void *stream)
{
    if (!d_slices || !d_state || n_slices > 65535 || max_n > 16384 || angle < 0)
        return (int)hipErrorInvalidValue;
    if (n_slices <= 0 || max_n <= 0 || max_d <= 0)
        return 0;
    hipStream_t st = (hipStream_t)stream;
    if (img_dtype == UIR_DTYPE_F32) {
        switch (dtype) {
        case UIR_DTYPE_F16: return launch_sart_line<__half, float>(d_slices, d_state, n_slices, max_n, max_d, angle, st);
        case UIR_DTYPE_BF16:
            return launch_sart_line<__hip_bfloat16, float>(d_slices, d_state, n_slices, max_n, max_d, angle, st);
        case UIR_DTYPE_F32: return launch_sart_line<float, float>(d_slices, d_state, n_slices, max_n, max_d, angle, st);
        default: return (int)hipErrorInvalidValue;
        }
    }
    if (img_dtype != dtype)
        return (int)hipErrorInvalidValue;
    switch (dtype) {
    case UIR_DTYPE_F16: return launch_sart_line<__half, __half>(d_slices, d_state, n_slices, max_n, max_d, angle, st);
    case UIR_DTYPE_BF16:
        return launch_sart_line<__hip_bfloat16, __hip_bfloat16>(d_slices, d_state, n_slices, max_n, max_d, angle, st);
    default: return (int)hipErrorInvalidValue;
    }
}


int uirk_fp_line(const float *vol, const float *r_sino, float *w_sino, const float *ray_w,
                 const uirk_line_angle *d_tab, int family, int n, int n_dets, int a0, int na, float corr,
                 int few_rays, void *stream)
{
    if (na <= 0 || n_dets <= 0 || n <= 0)
        return 0;
    if (na > 65535 || (family != 0 && family != 1))
        return (int)hipErrorInvalidValue;
    LineFpArgs a{vol, r_sino, w_sino, ray_w, d_tab, n, n_dets, a0, corr};
    if (family == 0 && !few_rays) {
        hipLaunchKernelGGL((k_fp_line<64, 0>), dim3((n_dets + 255) / 256, na), dim3(256), 0, (hipStream_t)stream, a);
    } else if (family == 0) {
        hipLaunchKernelGGL((k_fp_line<4, 0>), dim3((n_dets + 15) / 16, na), dim3(256), 0, (hipStream_t)stream, a);
    } else {
        hipLaunchKernelGGL((k_fp_line<4, 1>), dim3((n_dets + 15) / 16, na), dim3(256), 0, (hipStream_t)stream, a);
    }
    return (int)hipGetLastError();
}

int uirk_bp_line(const float *r_vol, float *w_vol, const float *sino, const float *pix_w, int posc,
                 const uirk_line_angle *d_tab, int n, int n_dets, int a0, int na, float relax, void *stream)
{
    if (n <= 0)
        return 0;
    if (na < 0 || n_dets <= 0)
        return (int)hipErrorInvalidValue;
    LineBpArgs a{r_vol, w_vol, sino, pix_w, d_tab, n, n_dets, a0, na, posc, relax};
    hipLaunchKernelGGL(k_bp_line, dim3((n + 63) / 64, (n + 15) / 16), dim3(256), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

} // extern "C"
