// uir_fast.hip -- table-driven, LDS-staged Joseph projector pair for gfx950.
//
// One kernel family serves two paths:
//   REF  = true : the reference's FP32 arithmetic operation for operation
//                 (exact incremental float ray stepping from the double start
//                 coordinate, 64-sample blocked sums, per-run BP sums, clamped
//                 detector coordinate), bit-identical to the one-thread-per-ray
//                 restatement of projector-parallel-joseph.cl.
//   REF  = false: the batched fast path.  FP32 interpolation and accumulation
//                 on FP32 / FP16 / BF16 storage, exact (double-anchored) ray
//                 coordinates, fused SIRT epilogues.  Its FP is k_fp_pairs
//                 (staged (e, d) pairs, one floor-convert and one FMA per
//                 sample) or, for geometries whose angle blocks need windows
//                 wider than the pairs' (sparse angles on large images),
//                 k_fp_tab<..., false, ...> (single staged values, fract).
// Work is described by tables (uirk_fp_item / uirk_bp_item) over any number
// of slices of any geometry; the kernels stride over the table (launched with
// one workgroup per item, so that the hardware's dispatch balances the items'
// uneven costs) and every item writes disjoint outputs, so results do not
// depend on the grid or on which other slices share the launch.
//
// FP: a workgroup owns 64 detectors x up to 4 * APT angles of one direction
//     family (the fast pair path: one angle and APT detectors per thread).
//     It walks the image in chunks of lines (rows for the vertical family,
//     columns for the horizontal one: 16 per chunk in the pair path, 32 in
//     the single-value one); each chunk's footprint is staged into LDS
//     line-major (transposed for the horizontal family), so both families
//     read their two taps as adjacent LDS words (one 8-byte pair in the pair
//     path).  Both projectors are bound by these LDS reads, 8 bytes per
//     sample.
// BP: a workgroup owns a 64 x 64 pixel tile (one column and 16 rows per
//     thread); sinogram rows and their sin/cos are staged 32 angles at a time
//     over the tile's detector footprint.
// Inner loops issue the LDS reads of all independent rays / rows of a step
// before consuming any, so LDS latency overlaps within the wave.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "uir_device.h"

namespace {

constexpr int kFpLines = UIRK_FP_LINES;
constexpr int kFpW = UIRK_FP_WMAX_SINGLE; // single-value staging (k_fp_tab)
constexpr int kFpLs = kFpW + 1;            // LDS line stride (odd: conflict-free transposed stores)
constexpr int kBpT = UIRK_BP_TILE;
constexpr int kBpWb = 100;                 // staged bins per angle (>= 64 * sqrt(2) + margin)

// Loads and stores through global-address-space pointers: the buffers come
// from descriptors in memory, so without the cast the compiler can only emit
// flat accesses (which also hold the LDS counter).
#define UIRK_GLOBAL __attribute__((address_space(1)))

// Offsets are 32-bit (every buffer holds < 2^32 elements) so that, with the
// base pointer uniform (see uniform()), loads use a scalar base plus a 32-bit
// vector offset.
// Every buffer is < 4 GiB (checked on the host), so byte offsets are 32-bit.
template <typename E> __device__ __forceinline__ const UIRK_GLOBAL E *gat(const void *p, uint32_t i)
{
    return (const UIRK_GLOBAL E *)((const UIRK_GLOBAL char *)p + i * (uint32_t)sizeof(E));
}
template <typename E> __device__ __forceinline__ UIRK_GLOBAL E *gat_w(void *p, uint32_t i)
{
    return (UIRK_GLOBAL E *)((UIRK_GLOBAL char *)p + i * (uint32_t)sizeof(E));
}
__device__ __forceinline__ float gload(const float *p, uint32_t i) { return *gat<float>(p, i); }
__device__ __forceinline__ float gload(const __half *p, uint32_t i)
{
    return __half2float(__ushort_as_half(*gat<unsigned short>(p, i)));
}
__device__ __forceinline__ float gload(const __hip_bfloat16 *p, uint32_t i)
{
    return __uint_as_float((unsigned)*gat<unsigned short>(p, i) << 16);
}
__device__ __forceinline__ void gstore(float *p, uint32_t i, float v) { *gat_w<float>(p, i) = v; }
__device__ __forceinline__ void gstore(__half *p, uint32_t i, float v)
{
    *gat_w<unsigned short>(p, i) = __half_as_ushort(__float2half(v));
}
__device__ __forceinline__ void gstore(__hip_bfloat16 *p, uint32_t i, float v)
{
    const __hip_bfloat16 h = __float2bfloat16(v);
    unsigned short bits;
    __builtin_memcpy(&bits, &h, sizeof bits);
    *gat_w<unsigned short>(p, i) = bits;
}

// A pointer that is the same for every lane (read from a descriptor), moved
// to scalar registers.
template <typename P> __device__ __forceinline__ P *uniform(P *p)
{
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (P *)(((unsigned long long)hi << 32) | lo);
}

// floor(x) converted to int in one instruction (no memory access).
__device__ __forceinline__ int cvt_flr(float x)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// round-to-nearest-even(x) for |x| < 2^22 with one float addition: x + 1.5 2^23
// lies in [2^23, 2^24), where floats are the integers, so its bits count x's
// rounded value up from those of 1.5 2^23.  (A float add issues at twice the
// rate of a float-to-int conversion.)  With x = w - 1/2 this is floor(w),
// except at integer w, where it may give w - 1: the pair FP reads
// piecewise-linear (e, d) entries, whose two segments agree at their shared
// end, so either index gives the same value there.
__device__ __forceinline__ int rne_index(float x) { return __float_as_int(x + 12582912.0f) - 0x4B400000; }

// -(float)n for 0 <= n < 2^23, assembled from integer operations: for a
// wave-uniform n the compiler keeps them, and the result, in scalar
// registers (a float conversion would put it in a vector register).
__device__ __forceinline__ float neg_float_of(int n)
{
    const int e = 31 - __builtin_clz((unsigned)n | 1u);
    const unsigned bits = 0x80000000u | ((unsigned)(127 + e) << 23) | (((unsigned)n << (23 - e)) & 0x7FFFFFu);
    return n == 0 ? -0.0f : __uint_as_float(bits);
}

// Entry rne_index(x) of an LDS table through a base biased by the bits of
// 1.5 2^23 (lds_rne_base, once per table): the sample's address is then one
// shift-add of x + 1.5 2^23's bits.  Unsigned 32-bit arithmetic wraps, so
// the biased base may lie outside LDS; the sum is the entry's address.
template <typename E> using lds_ptr = const __attribute__((address_space(3))) E *;
template <typename E> __device__ __forceinline__ uint32_t lds_rne_base(const E *table)
{
    return (uint32_t)(uintptr_t)(lds_ptr<E>)table - 0x4B400000u * (uint32_t)sizeof(E);
}
template <typename E> __device__ __forceinline__ E lds_rne_read(uint32_t base, float x)
{
    const uint32_t a = (uint32_t)__float_as_int(x + 12582912.0f) * (uint32_t)sizeof(E) + base;
    return *(const E *)(lds_ptr<E>)(uintptr_t)a;
}

// The line model's remapped fraction at coordinate x of the pair FP, whose
// entry is rne_index(x): f - 1/2 = x - rne(x) (exact), and
// phi = clamp((f - 1/2) k + b, 0, 1).
__device__ __forceinline__ float line_phi(float x, float k, float b)
{
    const float fh = x - ((x + 12582912.0f) - 12582912.0f);
    return __builtin_amdgcn_fmed3f(__builtin_fmaf(fh, k, b), 0.0f, 1.0f);
}

// x - floor(x); exact for x >= 0 (the callers guarantee it).
__device__ __forceinline__ float fract_pos(float x) { return __builtin_amdgcn_fractf(x); }

// Item range of this workgroup under an XCD-aware partition (see uirk_xcd_part).
struct ItemRange {
    int first, end, stride;
};

__device__ __forceinline__ ItemRange item_range(const uirk_xcd_part &part, int n_items)
{
    const int nx = part.n_xcd;
    if (nx > 1 && (gridDim.x % nx) == 0) {
        const int x = blockIdx.x % nx;
        return ItemRange{part.start[x] + (int)(blockIdx.x / nx), part.start[x + 1], (int)(gridDim.x / nx)};
    }
    return ItemRange{(int)blockIdx.x, n_items, (int)gridDim.x};
}

__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// A double held in lane `ln` (uniform) of every lane's register v.
__device__ __forceinline__ double readlane_f64(double v, int ln)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, ln);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), ln);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Workgroup-wide (4 waves) min of lo and max of hi.  Its first barrier also
// ends the previous chunk's reads of the staging area.
__device__ __forceinline__ void block_lohi(float &lo, float &hi, float (*red)[4], int lane, int wave)
{
    lo = wave_min(lo);
    hi = wave_max(hi);
    __syncthreads();
    if (lane == 0) {
        red[0][wave] = lo;
        red[1][wave] = hi;
    }
    __syncthreads();
    lo = fminf(fminf(red[0][0], red[0][1]), fminf(red[0][2], red[0][3]));
    hi = fmaxf(fmaxf(red[1][0], red[1][1]), fmaxf(red[1][2], red[1][3]));
}

// Stages one FP chunk: lines [l0, l0 + nl) (rows for the vertical family,
// columns for the horizontal one) x positions [pmin, pmin + W) go to
// lds[line * kFpLs + position - pmin], W <= kFpW.  Every thread issues a batch
// of loads (from clamped, always valid indices, no branches) before writing
// LDS; texels outside the image are then replaced by the zero border.
// Batches cover 64 window positions each, so only the footprint's positions
// are loaded.
template <typename TI>
__device__ __forceinline__ void fp_stage(float (&lds)[kFpLines * kFpLs], const TI *img, int N, bool hor, int l0, int nl,
                                         int pmin, int W)
{
    static_assert(kFpLines == 32 && kFpW % 64 == 0, "staging layout: 32 lines, 64-position batches");
    constexpr int kQ = 8;                  // loads per batch: 64 positions x 32 lines / 256 threads
    constexpr int kPer = kQ * (kFpW / 64); // loads per thread over the widest window
    if (hor) {
        // line = column l0 + l, position = row pmin + p; coalesced along columns
        const int l = threadIdx.x & 31;
        const int col = min(l0 + l, N - 1);
        for (int q0 = 0; q0 < kPer; q0 += kQ) {
            if (8 * q0 >= W)
                break;
            float v[kQ];
#pragma unroll
            for (int e = 0; e < kQ; ++e) {
                const int p = (threadIdx.x >> 5) + 8 * (q0 + e);
                const int row = pmin + p;
                const int rc = min(max(row, 0), N - 1);
                const float t = gload(img, (uint32_t)rc * (uint32_t)N + (uint32_t)col);
                v[e] = (row == rc && l < nl) ? t : 0.0f;
            }
#pragma unroll
            for (int e = 0; e < kQ; ++e) {
                const int p = (threadIdx.x >> 5) + 8 * (q0 + e);
                if (p < W)
                    lds[l * kFpLs + p] = v[e];
            }
        }
    } else {
        // line = row l0 + l, position = column pmin + p: batch q0 holds
        // positions [64 (q0 / kQ), + 64) of all 32 lines
        for (int q0 = 0; q0 < kPer; q0 += kQ) {
            const int pb = q0 / kQ;
            if (64 * pb >= W)
                break;
            float v[kQ];
#pragma unroll
            for (int e = 0; e < kQ; ++e) {
                const int l = (threadIdx.x >> 6) + 4 * e;
                const int p = (threadIdx.x & 63) + 64 * pb;
                const int col = pmin + p;
                const int cc = min(max(col, 0), N - 1);
                const int row = min(l0 + l, N - 1);
                const float t = gload(img, (uint32_t)row * (uint32_t)N + (uint32_t)cc);
                v[e] = (col == cc && l < nl) ? t : 0.0f;
            }
#pragma unroll
            for (int e = 0; e < kQ; ++e) {
                const int l = (threadIdx.x >> 6) + 4 * e;
                const int p = (threadIdx.x & 63) + 64 * pb;
                if (p < W)
                    lds[l * kFpLs + p] = v[e];
            }
        }
    }
}

// Buffer resources for the fast FP's staging.  A raw buffer load whose byte
// offset is at or beyond the resource's size returns zero and touches no
// memory, which gives the zero border outside the image without clamping or
// selecting: negative indices become offsets above 2^32 - 2^24, beyond any
// resource here (the batch keeps N <= 16384, so an image is at most 1 GiB).
// The immediate and scalar offsets of the loads are always zero, so the range
// check sees the whole offset.
constexpr int kBufDword3 = 0x00020000; // gfx9 raw-buffer format word (32-bit data, no swizzle)

template <typename TI> __device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const TI *base, uint32_t elems)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)(elems * (uint32_t)sizeof(TI)), kBufDword3);
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, const float *, int i)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (uint32_t)i * 4u, 0, 0));
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, const __half *, int i)
{
    return __half2float(__ushort_as_half(__builtin_amdgcn_raw_buffer_load_b16(r, (uint32_t)i * 2u, 0, 0)));
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, const __hip_bfloat16 *, int i)
{
    return __uint_as_float((unsigned)__builtin_amdgcn_raw_buffer_load_b16(r, (uint32_t)i * 2u, 0, 0) << 16);
}

// Fast FP staging as (e, d) pairs.  For line l and window position p (texel
// t_p at position pmin + p, zero outside the image):
//   d = t_{p+1} - t_p,  e = t_p - (p - kFpPH - 1/2) d,
// so the interpolated value at window coordinate w in [p, p + 1) is
// e + (w - kFpPH - 1/2) d: one FMA per sample, no fract and no second tap,
// and the entry index is rne_index(w - kFpPH - 1/2), one float add.  The
// origin kFpPH is the window's middle, which keeps |w - kFpPH| <= kFpPW / 2.
// PLAIN (the line model) stages e = t_p: its samples remap the fraction.
constexpr int kFpPW = UIRK_FP_WMAX; // widest pair window
constexpr int kFpPL = 16;           // lines per pair-staged chunk (25 KB of LDS: 6 workgroups per CU)
constexpr int kFpPs = kFpPW + 1;    // pair line stride (odd)
constexpr int kFpPH = kFpPW / 2;

// A staged (e, d) pair, or (e, d, e', d') for a slice pair (the BP's mode 3).
template <bool PAIR> struct fp_entry {
    using type = float2;
};
template <> struct fp_entry<true> {
    using type = float4;
};
template <bool PAIR> __device__ __forceinline__ typename fp_entry<PAIR>::type make_entry(float e, float d, float e2,
                                                                                         float d2);
template <> __device__ __forceinline__ float2 make_entry<false>(float e, float d, float, float)
{
    return make_float2(e, d);
}
template <> __device__ __forceinline__ float4 make_entry<true>(float e, float d, float e2, float d2)
{
    return make_float4(e, d, e2, d2);
}

template <typename TI, bool PLAIN>
__device__ __forceinline__ void fp_stage_pairs(float2 (&lds)[kFpPL * kFpPs], const TI *img, int N, bool hor, int l0,
                                               int nl, int pmin, int W)
{
    static_assert(kFpPL % 4 == 0 && kFpPL <= 32 && kFpPW % 64 == 0 && kFpPs > kFpPW, "staging layout");
    // Stores are unconditional: positions up to the end of the last 64-wide
    // batch (< kFpPW) and lines below kFpPL lie inside the staging area, and
    // entries at positions >= W or lines >= nl are never sampled.  (Guarding
    // them would let the compiler sink each load into its store's branch and
    // wait for it there.)
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (!hor) {
        // line = row l0 + l, position = column pmin + p.  A wave instruction
        // covers 64 consecutive positions of one row; a thread loads its
        // position in the 8 rows wave, wave + 4, ... before using any.  One
        // resource per row: columns outside [0, N) read as zero; rows
        // l >= nl get an empty resource.
        for (int pb = 0; pb < W; pb += 64) {
            const int p = pb + lane;
            const int c = pmin + p;
            float t0[kFpPL / 4], t1[kFpPL / 4];
#pragma unroll
            for (int i = 0; i < kFpPL / 4; ++i) {
                const int l = wave + 4 * i;
                const __amdgpu_buffer_rsrc_t r = buf_rsrc(img + (uint32_t)min(l0 + l, N - 1) * (uint32_t)N,
                                                          l < nl ? (uint32_t)N : 0u);
                t0[i] = bload(r, img, c);
                t1[i] = bload(r, img, c + 1);
            }
            const float h = (float)(p - kFpPH) - 0.5f; // entries for coordinates w - 1/2
#pragma unroll
            for (int i = 0; i < kFpPL / 4; ++i) {
                const float d = t1[i] - t0[i];
                lds[(wave + 4 * i) * kFpPs + p] = make_float2(PLAIN ? t0[i] : __builtin_fmaf(-h, d, t0[i]), d);
            }
        }
    } else {
        // line = column l0 + l (l = lane mod kFpPL), position = row pmin + p.
        // A wave instruction covers kFpPL consecutive columns of 64 / kFpPL
        // rows (coalesced per row), every lane busy; a thread loads kPer + 1
        // consecutive rows for kPer pairs.  One resource over the image: rows
        // outside [0, N) read as zero; columns at or beyond N occur only for
        // lines l >= nl.
        static_assert(64 % kFpPL == 0, "lanes split evenly over the staged lines");
        constexpr int kSub = 64 / kFpPL;  // row groups per wave instruction
        constexpr int kPer = 16 / kSub;   // pairs per thread per 16 positions of a wave
        const int l = lane % kFpPL;
        const int col = l0 + l;
        const __amdgpu_buffer_rsrc_t r = buf_rsrc(img, (uint32_t)N * (uint32_t)N);
        for (int pb = 0; pb < W; pb += 64) {
            if (pb + 16 * wave >= W)
                break;
            const int p0 = pb + 16 * wave + kPer * (lane / kFpPL);
            const int i0 = (pmin + p0) * N + col;
            float t[kPer + 1];
#pragma unroll
            for (int k = 0; k <= kPer; ++k)
                t[k] = bload(r, img, i0 + k * N);
            const float h0 = (float)(p0 - kFpPH) - 0.5f; // entries for coordinates w - 1/2
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const float d = t[k + 1] - t[k];
                lds[l * kFpPs + p0 + k] = make_float2(PLAIN ? t[k] : __builtin_fmaf(-(h0 + (float)k), d, t[k]), d);
            }
        }
    }
}

// Start coordinate (line 0) and per-line step of the fast path's ray of
// detector coordinate tdet at angle ang, in double (projector-parallel-joseph.cl:51-73
// with exact arithmetic).  Every lane that evaluates the same ray gets the
// same bits.
__device__ __forceinline__ void fast_ray(const float *trig, int ang, bool hor, double tdet, int N, double &p0,
                                         double &sd)
{
    const float st = gload(trig, 2 * ang), ct = gload(trig, 2 * ang + 1);
    const double det_step = hor ? -1.0 / (double)st : 1.0 / (double)ct;
    sd = hor ? (double)ct / (double)st : (double)st / (double)ct;
    p0 = tdet * det_step - 0.5 * (double)N * sd + 0.5 * (double)N;
}

// ------------------------------------------------------------------------
// Forward projection
// ------------------------------------------------------------------------
// FP over a work table with single-value staging: the reference-compatible
// FP (REF) and the fast path for geometries whose angle blocks need windows
// wider than the pair staging holds (fp_single tiles).
template <typename T, typename TI, bool REF, int APT>
__global__ __launch_bounds__(256) void k_fp_tab(const uirk_slice *__restrict__ slices, uirk_slice one,
                                                const uirk_fp_item *__restrict__ items, int n_items,
                                                uirk_xcd_part part, int *__restrict__ err)
{
    __shared__ float lds[kFpLines * kFpLs];
    __shared__ float red[2][4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const ItemRange range = item_range(part, n_items);

    for (int it = range.first; it < range.end; it += range.stride) {
        const uirk_fp_item item = items[it];
        if (item.na > 4 * APT) { // planned for a wider variant: report, never drop angles silently
            if (threadIdx.x == 0)
                atomicOr(err, 4);
            continue;
        }
        const uirk_slice *S = slices ? &slices[item.slice] : &one;
        const int N = S->n, D = S->n_dets;
        const TI *img = uniform((const TI *)S->x);
        const float *trig = uniform(S->trig);
        const int s = item.d0 + lane;
        const bool hor = item.dir == 0;

        // projector-parallel-joseph.cl:51-66 (detector quantities use n_dets)
        const float required_width = S->axis * 2;
        const float diff = (float)D - required_width;
        const float rot = diff / 2.0f;
        const float origin_shift = (float)N / 2.0f;
        const bool in_range = s < D;
        const bool det_on = in_range && !((diff < 0 && (float)s < fabsf(diff)) || (diff > 0 && (float)s > required_width));
        const double tdet = ((0.5 + (double)rot) + (double)s) - (double)(0.5f * (float)D);

        float pos[APT], stp[APT], acc[APT], inner[APT];
        double p0d[APT], sd[APT];
        bool on[APT];
#pragma unroll
        for (int j = 0; j < APT; ++j) {
            const int k = wave * APT + j;
            on[j] = det_on && k < item.na;
            const int ang = item.a0 + (k < item.na ? k : 0);
            const float st = gload(trig, 2 * ang), ct = gload(trig, 2 * ang + 1);
            double p0;
            if (REF) {
                float det_step, step;
                if (hor) {
                    det_step = -1.0f / st;
                    step = ct / st;
                } else {
                    det_step = 1.0f / ct;
                    step = st / ct;
                }
                // start coordinate: double expression rounded to float (:73, :134)
                p0 = tdet * (double)det_step + (double)((-origin_shift) * step) + (double)(0.5f * (float)N);
                sd[j] = (double)step;
            } else {
                // exact coordinates: steps and chunk anchors in double
                const double det_step = hor ? -1.0 / (double)st : 1.0 / (double)ct;
                sd[j] = hor ? (double)ct / (double)st : (double)st / (double)ct;
                p0 = tdet * det_step - 0.5 * (double)N * sd[j] + 0.5 * (double)N;
            }
            p0d[j] = p0;
            pos[j] = (float)p0;
            stp[j] = (float)sd[j];
            acc[j] = 0.0f;
            inner[j] = 0.0f;
        }

        // Lines before l_lo read only the zero border: the reference-compatible
        // path replays their float steps exactly, the fast path re-anchors.
        if (REF) {
            for (int l = 0; l < item.l_lo; ++l) {
#pragma unroll
                for (int j = 0; j < APT; ++j)
                    pos[j] += stp[j];
            }
        }
        for (int l0 = item.l_lo; l0 < item.l_hi; l0 += kFpLines) {
            const int nl = min(kFpLines, item.l_hi - l0);
            // Footprint of this block's rays over the chunk.
            float lo = 3.0e38f, hi = -3.0e38f;
#pragma unroll
            for (int j = 0; j < APT; ++j) {
                if (!REF)
                    pos[j] = (float)(p0d[j] + (double)l0 * sd[j]);
                if (on[j]) {
                    const float a = pos[j], b = pos[j] + (float)(nl - 1) * stp[j];
                    lo = fminf(lo, fminf(a, b));
                    hi = fmaxf(hi, fmaxf(a, b));
                }
            }
            block_lohi(lo, hi, red, lane, wave);
            if (!(lo <= hi))
                break; // no active ray in this workgroup
            const int pmin = (int)floorf(lo - 0.5f) - 1;
            const int W = (int)floorf(hi - 0.5f) + 3 - pmin;
            // A footprint wider than the staging window is reported and the
            // chunk's samples are skipped (the host sizes angle blocks so that
            // this does not happen); no sample ever reads outside the window.
            const bool overflow = W > kFpW;
            if (overflow && threadIdx.x == 0)
                atomicOr(err, 1);
            const bool outside = overflow || pmin + W <= 0 || pmin >= N;
            if (!outside)
                fp_stage(lds, img, N, hor, l0, nl, pmin, W);
            __syncthreads();

            // Inactive rays sample a fixed in-window position with zero step;
            // their sums are never written.
            const float park = lo + 1.0f;
            if (REF) {
                if (!outside) {
                    float pp[APT], se[APT];
#pragma unroll
                    for (int j = 0; j < APT; ++j) {
                        pp[j] = on[j] ? pos[j] : park;
                        se[j] = on[j] ? stp[j] : 0.0f;
                    }
                    for (int l = 0; l < nl; ++l) {
                        const float *L = lds + l * kFpLs - pmin;
                        float t0[APT], t1[APT], b[APT];
#pragma unroll
                        for (int j = 0; j < APT; ++j) {
                            const float v = pp[j] - 0.5f;
                            const float fv = floorf(v);
                            const int i0 = (int)fv;
                            b[j] = v - fv;
                            t0[j] = L[i0];
                            t1[j] = L[i0 + 1];
                        }
#pragma unroll
                        for (int j = 0; j < APT; ++j) {
                            const float omb = 1.0f - b[j];
                            inner[j] += omb * t0[j] + b[j] * t1[j];
                            pp[j] += se[j];
                        }
                        if (((l0 + l + 1) & 63) == 0) { // 64-sample blocks (:82-90)
#pragma unroll
                            for (int j = 0; j < APT; ++j) {
                                acc[j] += inner[j];
                                inner[j] = 0.0f;
                            }
                        }
                    }
#pragma unroll
                    for (int j = 0; j < APT; ++j)
                        if (on[j])
                            pos[j] = pp[j];
                } else {
                    // outside the image: the samples are zero, the exact
                    // float stepping of the coordinates continues
                    for (int l = 0; l < nl; ++l) {
#pragma unroll
                        for (int j = 0; j < APT; ++j)
                            pos[j] += stp[j];
                        if (((l0 + l + 1) & 63) == 0) {
#pragma unroll
                            for (int j = 0; j < APT; ++j) {
                                acc[j] += inner[j];
                                inner[j] = 0.0f;
                            }
                        }
                    }
                }
            } else if (!outside) {
                const float off = 0.5f + (float)pmin;
                float q[APT], st[APT];
#pragma unroll
                for (int j = 0; j < APT; ++j) {
                    q[j] = on[j] ? pos[j] - off : park - off;
                    st[j] = on[j] ? stp[j] : 0.0f;
                }
                // Two lines per step: the LDS reads of 2 x APT samples are
                // in flight before the first is consumed.
                int l = 0;
                for (; l + 2 <= nl; l += 2) {
                    const float *L0 = lds + l * kFpLs, *L1 = L0 + kFpLs;
                    float t0[2][APT], t1[2][APT], a[2][APT];
#pragma unroll
                    for (int j = 0; j < APT; ++j) {
                        const float q1 = q[j] + st[j];
                        const int i0 = cvt_flr(q[j]), i1 = cvt_flr(q1);
                        a[0][j] = fract_pos(q[j]);
                        a[1][j] = fract_pos(q1);
                        t0[0][j] = L0[i0];
                        t1[0][j] = L0[i0 + 1];
                        t0[1][j] = L1[i1];
                        t1[1][j] = L1[i1 + 1];
                        q[j] = q1 + st[j];
                    }
#pragma unroll
                    for (int j = 0; j < APT; ++j) {
                        acc[j] = __builtin_fmaf(a[0][j], t1[0][j] - t0[0][j], acc[j] + t0[0][j]);
                        acc[j] = __builtin_fmaf(a[1][j], t1[1][j] - t0[1][j], acc[j] + t0[1][j]);
                    }
                }
                if (l < nl) {
                    const float *L = lds + l * kFpLs;
#pragma unroll
                    for (int j = 0; j < APT; ++j) {
                        const int i0 = cvt_flr(q[j]);
                        const float a = fract_pos(q[j]);
                        const float t0 = L[i0], t1 = L[i0 + 1];
                        acc[j] = __builtin_fmaf(a, t1 - t0, acc[j] + t0);
                        q[j] += st[j];
                    }
                }
            }
        }

        const T *bsrc = uniform((const T *)S->b);
        T *rdst = uniform((T *)S->r);
        const float *ray_w = uniform(S->ray_w);
        const float corr = S->corr;
#pragma unroll
        for (int j = 0; j < APT; ++j) {
            const int k = wave * APT + j;
            if (!in_range || k >= item.na)
                continue;
            const int ang = item.a0 + k;
            const uint32_t o = (uint32_t)ang * (uint32_t)D + (uint32_t)s;
            const float bval = gload(bsrc, o);
            float w;
            if (on[j]) {
                const float sum = REF ? acc[j] + inner[j] : acc[j];
                w = bval + sum * corr; // :92-94
            } else {
                w = bval; // inactive detector: the reference leaves the value as read
            }
            if (ray_w)
                w = w * gload(ray_w, o);
            gstore(rdst, o, w);
        }
    }
}

// Fast-path FP: the work table of k_fp_tab, exact (double-anchored) ray
// coordinates per chunk, the chunk's footprint staged as (e, d) pairs.
//  - The footprint is the hull of the extreme rays: for one angle a ray's
//    chunk positions are monotone in the detector index, so the first and
//    last active detectors bound it.  Each wave evaluates those angles x 2
//    detectors x 2 chunk ends itself (no workgroup reduction, no barrier),
//    with the same operations as the ray lanes, for 64 chunks at a time:
//    one chunk per lane.
//  - A sample costs an add (coordinate), a floor-convert, an address, one
//    8-byte LDS read and two FMA-class operations.
//  - LINE: the line model (exact chord lengths).  Along a ray the chords of
//    one line of pixels sum to 1/M, so a sample is the lerp e + phi(f) d of
//    the plain pair (t_p, t_{p+1} - t_p) with the fraction f remapped by
//    phi(f) = clamp((f - 1/2) fp_k + fp_b, 0, 1) (a ramp of width m / M about
//    1/2, or a step with the tie rule for axis-parallel rays), times 1/M.
//    Positions lie half a line step further along than joseph's (the
//    model's exact geometry), and every ray is active.  Four more
//    operations per sample.  The remap multiplies position errors by M / m,
//    so a chunk's start coordinate is rounded once from double, relative to
//    the window, and each line's coordinate is one FMA from it (joseph steps
//    by float additions).
template <typename T, typename TI, int APT, bool LINE>
__global__ __launch_bounds__(256, APT == 8 ? 4 : 6) void k_fp_pairs(const uirk_slice *__restrict__ slices, uirk_slice one,
                                                  const uirk_fp_item *__restrict__ items, int n_items,
                                                  uirk_xcd_part part, int *__restrict__ err)
{
    __shared__ float2 lds[kFpPL * kFpPs];
    constexpr int kA = 4 * APT;               // angles per item
    constexpr int kC = (4 * kA + 63) / 64;    // footprint corners per lane
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const ItemRange range = item_range(part, n_items);

    for (int it = range.first; it < range.end; it += range.stride) {
        const uirk_fp_item item = items[it];
        const uirk_slice *S = slices ? &slices[item.slice] : &one;
        const int N = S->n, D = S->n_dets;
        if (item.na > kA || N > 16384) { // planned for a wider variant / beyond the staging's offsets
            if (threadIdx.x == 0)
                atomicOr(err, 4);
            continue;
        }
        const TI *img = uniform((const TI *)S->x);
        const float *trig = uniform(S->trig);
        const int s = item.d0 + lane;
        const bool hor = item.dir == 0;

        // projector-parallel-joseph.cl:51-66 (detector quantities use n_dets)
        const float required_width = S->axis * 2;
        const float diff = (float)D - required_width;
        const float rot = diff / 2.0f;
        const bool in_range = s < D;
        const bool det_on =
            in_range && (LINE || !((diff < 0 && (float)s < fabsf(diff)) || (diff > 0 && (float)s > required_width)));
        const double tbase = 0.5 + (double)rot;
        const double thalf = (double)(0.5f * (float)D);

        // Rays of this thread: one angle, APT detectors kDL apart.  A wave
        // instruction (fixed j) then covers kDL adjacent detectors of each of
        // the wave's APT angles, whose positions lie closer together than those
        // of 64 adjacent detectors of one angle (fewer distinct LDS pairs per
        // read), and a thread divides once per item instead of APT times.
        constexpr int kDL = 64 / APT;
        const int k_ang = wave * APT + lane / kDL; // angle within the item
        const bool ang_on = k_ang < item.na;
        float stp[APT], acc[APT];
        double p0d[APT], sd[APT];
        bool on[APT], ray_in[APT];
        float lk = 0.0f, lb = 0.0f, lhm = 1.0f; // LINE: the angle's phi and 1/M
        {
            // fast_ray, with the angle's quantities computed once
            const int ang = item.a0 + (ang_on ? k_ang : 0);
            const float st = gload(trig, 2 * ang), ct = gload(trig, 2 * ang + 1);
            const double det_step = hor ? -1.0 / (double)st : 1.0 / (double)ct;
            const double sdv = hor ? (double)ct / (double)st : (double)st / (double)ct;
            if constexpr (LINE) {
                const float *lt = uniform(S->line);
                lk = gload(lt, 8 * ang);
                lb = gload(lt, 8 * ang + 1);
                lhm = gload(lt, 8 * ang + 2);
            }
#pragma unroll
            for (int j = 0; j < APT; ++j) {
                const int sj = item.d0 + lane % kDL + kDL * j;
                ray_in[j] = sj < D;
                const bool dj_on = ray_in[j] && (LINE || !((diff < 0 && (float)sj < fabsf(diff)) ||
                                                           (diff > 0 && (float)sj > required_width)));
                on[j] = dj_on && ang_on;
                p0d[j] = ((tbase + (double)sj) - thalf) * det_step - 0.5 * (double)N * sdv + 0.5 * (double)N +
                         (LINE ? 0.5 * sdv : 0.0);
                sd[j] = sdv;
                stp[j] = (float)sdv;
                acc[j] = 0.0f;
            }
        }
        // footprint corners: (angle, first / last active detector, chunk start / end)
        const unsigned long long m = __ballot(det_on);
        double cp[kC], cs[kC];
        bool con[kC], cend[kC];
        if (m != 0) {
            const int s_first = item.d0 + __builtin_ctzll(m), s_last = item.d0 + 63 - __builtin_clzll(m);
#pragma unroll
            for (int e = 0; e < kC; ++e) {
                const int idx = lane + 64 * e;
                const int k = idx >> 2;
                con[e] = k < item.na;
                cend[e] = idx & 1;
                const int sx = (idx & 2) ? s_last : s_first;
                fast_ray(trig, item.a0 + (con[e] ? k : 0), hor, (tbase + (double)sx) - thalf, N, cp[e], cs[e]);
                if (LINE)
                    cp[e] += 0.5 * cs[e];
            }
        }

        // Chunk windows, 64 chunks at a time: lane c evaluates chunk cg + c
        // at every corner (read from the corner lanes), and each chunk then
        // reads its lo / hi from that lane (uniform).  The corners' (angle,
        // detector) pairs sit in lanes 4 k + 2 d (+ 64 per extra corner
        // register), the same values the per-lane corners hold.
        float c_lo = 0.0f, c_hi = 0.0f;
        int ci = 0;
        for (int l0 = item.l_lo; m != 0 && l0 < item.l_hi; l0 += kFpPL, ++ci) {
            const int nl = min(kFpPL, item.l_hi - l0);
            if ((ci & 63) == 0) {
                const int cl0 = l0 + lane * kFpPL; // this lane's chunk (beyond l_hi: unused)
                const double cl0d = (double)cl0;
                const float nm1 = (float)(min(kFpPL, item.l_hi - cl0) - 1);
                float lo_c = 3.0e38f, hi_c = -3.0e38f;
#pragma unroll
                for (int k = 0; k < kA; ++k) {
                    if (k >= item.na) // uniform
                        continue;
#pragma unroll
                    for (int dsel = 0; dsel < 2; ++dsel) {
                        const int idx = 4 * k + 2 * dsel;
                        const double cpe = readlane_f64(cp[idx >> 6], idx & 63);
                        const double cse = readlane_f64(cs[idx >> 6], idx & 63);
                        const float av = (float)(cpe + cl0d * cse);
                        const float bv = av + nm1 * (float)cse;
                        lo_c = fminf(lo_c, fminf(av, bv));
                        hi_c = fmaxf(hi_c, fmaxf(av, bv));
                    }
                }
                c_lo = lo_c;
                c_hi = hi_c;
            }
            const float lo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c_lo), ci & 63));
            const float hi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c_hi), ci & 63));
            const int pmin = (int)floorf(lo - 0.5f) - 1;
            const int W = (int)floorf(hi - 0.5f) + 3 - pmin;
            // A footprint wider than the staging window is reported and the
            // chunk's samples are skipped (the host sizes angle blocks so that
            // this does not happen); no sample ever reads outside the window.
            const bool overflow = W > kFpPW;
            if (overflow && threadIdx.x == 0)
                atomicOr(err, 1);
            const bool outside = overflow || pmin + W <= 0 || pmin >= N;
            if (outside)
                continue;
            __syncthreads(); // the previous chunk's reads are done
            fp_stage_pairs<TI, LINE>(lds, img, N, hor, l0, nl, pmin, W);
            __syncthreads();
            if (wave * APT >= item.na)
                continue; // this wave has no angle of the item: it only stages

            // Window coordinate relative to kFpPH, less 1/2 (rne_index of it is
            // the sample's entry); inactive rays sample a fixed in-window
            // position with zero step and are never written.
            const float off = 1.0f + (float)pmin + (float)kFpPH;
            const float park = lo + 1.0f - off;
            float u[APT], st[APT];
#pragma unroll
            for (int j = 0; j < APT; ++j) {
                if (LINE) {
                    u[j] = on[j] ? (float)(p0d[j] + (double)l0 * sd[j] - (double)off) : park;
                } else {
                    const float pos = (float)(p0d[j] + (double)l0 * sd[j]);
                    u[j] = on[j] ? pos - off : park;
                }
                st[j] = on[j] ? stp[j] : 0.0f;
            }
            const float2 *P = lds + kFpPH;
            int l = 0;
            // kQ lines per step (8 or 16 samples per thread), then single lines:
            // the LDS reads of a step are all in flight before the first is
            // consumed.  The coordinates advance by the same float additions
            // whatever the grouping.
            constexpr int kQ = APT >= 4 ? 2 : 8; // APT 4 fits 80 VGPRs (6 waves per SIMD) with 2
            for (; l + kQ <= nl; l += kQ) {
                const float2 *P0 = P + l * kFpPs;
                float2 v[kQ][APT];
                float w[kQ][APT];
#pragma unroll
                for (int j = 0; j < APT; ++j) {
                    w[0][j] = LINE ? __builtin_fmaf((float)l, st[j], u[j]) : u[j];
#pragma unroll
                    for (int q = 1; q < kQ; ++q)
                        w[q][j] = LINE ? __builtin_fmaf((float)(l + q), st[j], u[j]) : w[q - 1][j] + st[j];
#pragma unroll
                    for (int q = 0; q < kQ; ++q)
                        v[q][j] = P0[q * kFpPs + rne_index(w[q][j])];
                }
#pragma unroll
                for (int j = 0; j < APT; ++j) {
#pragma unroll
                    for (int q = 0; q < kQ; ++q)
                        acc[j] = __builtin_fmaf(LINE ? line_phi(w[q][j], lk, lb) : w[q][j], v[q][j].y,
                                                acc[j] + v[q][j].x);
                    if (!LINE) // LINE: u stays the chunk's start
                        u[j] = w[kQ - 1][j] + st[j];
                }
            }
            for (; l < nl; ++l) {
                const float2 *P0 = P + l * kFpPs;
#pragma unroll
                for (int j = 0; j < APT; ++j) {
                    const float uj = LINE ? __builtin_fmaf((float)l, st[j], u[j]) : u[j];
                    const float2 v = P0[rne_index(uj)];
                    acc[j] = __builtin_fmaf(LINE ? line_phi(uj, lk, lb) : uj, v.y, acc[j] + v.x);
                    if (!LINE)
                        u[j] += st[j];
                }
            }
        }

        const T *bsrc = uniform((const T *)S->b);
        T *rdst = uniform((T *)S->r);
        const float *ray_w = uniform(S->ray_w);
        const float corr = S->corr;
#pragma unroll
        for (int j = 0; j < APT; ++j) {
            if (!ray_in[j] || !ang_on)
                continue;
            const uint32_t o = (uint32_t)(item.a0 + k_ang) * (uint32_t)D + (uint32_t)(item.d0 + lane % kDL + kDL * j);
            // :92-94; an inactive detector keeps the value as read
            float w = gload(bsrc, o);
            if (on[j])
                w = w + (LINE ? acc[j] * lhm : acc[j]) * corr;
            if (ray_w)
                w = w * gload(ray_w, o);
            gstore(rdst, o, w);
        }
    }
}

// ------------------------------------------------------------------------
// Backprojection
// ------------------------------------------------------------------------
// A workgroup backprojects a tile of 64 x 4 ROWS pixels (one column per lane,
// rows wave + 4 q), NA angles at a time staged in LDS.  The window and, on the
// fast path, the sample coordinates belong to the 64 x 64 super-tile holding
// the tile, and each pixel sums its angles in order, so the result of a pixel
// does not depend on ROWS, NA or slice pairing.
// LINE (fast path only): the line model.  Bins outside the detector read 0
// (every ray is active, no clamp), bins are staged as single values (a
// sample reads two adjacent words in one instruction), and a sample weighs
// them by the chord lengths of the pixel on the two rays,
// w0 = clamp(bp_b0 - bp_k (f - 1/2), 0, 1/M), w1 = clamp(bp_b1 + bp_k (f - 1/2),
// 0, 1/M), f the fraction of the detector coordinate; pix_w, when given,
// multiplies the update (the model's pixel weights are not uniform).
// PAIR (fast path only): an item with item.pair set also backprojects slice
// item.slice + 1 of the same geometry; both slices' (e, d) pairs share a
// 16-byte entry, so a pixel-angle's coordinate, address and one LDS read
// serve both (the same operations per slice as alone).  With LINE the staged
// bin holds both slices' values (t, t'), and the chord weights are shared.
template <typename T, typename TI, bool REF, int NA, int ROWS, bool LINE, bool PAIR>
__global__ __launch_bounds__(256, REF ? 1 : (LINE ? 4 : (PAIR ? (NA == 32 ? 3 : 4) : 6))) void k_bp_tab(
    const uirk_slice *__restrict__ slices, uirk_slice one, const uirk_bp_item *__restrict__ items, int n_items,
    uirk_xcd_part part, int *__restrict__ err)
{
    static_assert(!PAIR || !REF, "slice pairs: the fast path only");
    using E = typename fp_entry<PAIR>::type;
    // reference-compatible path and line model: plain staged bins (line pairs:
    // both slices' values); fast path: (e, d) pairs
    __shared__ float lds[(REF || (LINE && !PAIR)) ? NA * kBpWb : 1];
    __shared__ float2 ldl[(LINE && PAIR) ? NA * kBpWb : 1];
    __shared__ E ldp[(REF || LINE) ? 1 : NA * kBpWb];
    __shared__ int s_umin[NA];
    __shared__ float s_trig[2 * NA];
    __shared__ float4 s_line[LINE ? NA : 1];
    __shared__ float s_base[REF ? 1 : NA]; // fast path: window coordinate at the super-tile's corner
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const ItemRange range = item_range(part, n_items);

    for (int it = range.first; it < range.end; it += range.stride) {
        const uirk_bp_item item = items[it];
        const uirk_slice *S = slices ? &slices[item.slice] : &one;
        const int N = S->n, D = S->n_dets;
        const bool paired = PAIR && item.pair != 0 && slices;
        const uirk_slice *S2 = paired ? &slices[item.slice + 1] : S;
        const T *sino = uniform((const T *)S->b);
        const T *sino2 = PAIR ? uniform((const T *)S2->b) : sino;
        const float *trig = uniform(S->trig);
        const float relax = S->relax;
        const int n_runs = S->n_runs;
        const int px = item.x0 + lane;

        // projector-parallel-joseph.cl:174-190 (detector quantities use n_dets)
        const float axis = S->axis;
        const float required_width = axis * 2;
        const float diff = (float)D - required_width;
        const float half_active = diff < 0 ? (float)D - axis : axis;
        const float edge0 = axis - half_active + 0.5f;
        const float edge1 = axis + half_active - 0.5f;
        const float rot = diff / 2.0f;
        const float det_shift = (float)D / 2.0f;
        const float origin_shift = (float)N / 2.0f;
        const float shift = det_shift - rot;
        const float fX = (float)px + 0.5f - origin_shift;
        // active bins for the fast path's clamp-to-edge staging, kept inside
        // [0, D - 1]: an axis near 0 or past D leaves the active window empty
        // or outside the detector, and every staging load must stay in the row
        const int bin_lo = min(max((int)floorf(edge0 - 0.5f), 0), D - 1);
        const int bin_hi = min(max((int)floorf(edge1 - 0.5f), 0), D - 1);
        // Clamping the bins equals the reference's clamp of u (:203-210) when
        // both edges sit on bin centres (edge - 0.5 an integer: a centred axis,
        // or one shifted by a multiple of 1/2).  Otherwise an edge falls inside
        // a bin, and the fast path clamps the coordinate itself (one more
        // operation per sample; uniform over the slice).
        const bool clamp_u =
            !REF && !LINE && ((edge0 - 0.5f) != floorf(edge0 - 0.5f) || (edge1 - 0.5f) != floorf(edge1 - 0.5f));
        // Window corners (pixel centres): those of the 64 x 64 super-tile that
        // holds the tile.  A tile of 32 rows stages its super-tile's window and,
        // on the fast path, computes its super-tile's sample coordinates, so a
        // pixel's result does not depend on the tile's rows or staged angles.
        const int y0s = item.y0 & ~(kBpT - 1);
        const float cx0 = (float)item.x0 + 0.5f - origin_shift;
        const float cx1 = cx0 + (float)(kBpT - 1);
        const float cy0 = (float)y0s + 0.5f - origin_shift;
        const float cy1 = cy0 + (float)(kBpT - 1);
        const float fY0 = (float)(item.y0 + wave) + 0.5f - origin_shift;

        float acc[ROWS], val[ROWS], acc2[PAIR ? ROWS : 1];
#pragma unroll
        for (int k = 0; k < ROWS; ++k) {
            acc[k] = 0.0f;
            val[k] = 0.0f;
            if constexpr (PAIR)
                acc2[k] = 0.0f;
        }
        // fast path: minus the offset of each of this thread's rows within the
        // super-tile, uniform over a wave (scalar registers)
        float nys[REF ? 1 : ROWS];
        if constexpr (!REF) {
            const int ys0 = item.y0 - y0s + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
            for (int q = 0; q < ROWS; ++q)
                nys[q] = neg_float_of(ys0 + 4 * q);
        }
        if (REF && !S->pix_w) {
            const TI *xin = uniform((const TI *)S->x);
#pragma unroll
            for (int k = 0; k < ROWS; ++k) {
                const int py = item.y0 + wave + 4 * k;
                acc[k] = (px < N && py < N) ? gload(xin, (uint32_t)py * (uint32_t)N + (uint32_t)px) : 0.0f;
            }
        }

        for (int r = 0; r < n_runs; ++r) {
            const int roff = (int)S->run_off[r], rn = (int)S->run_n[r];
            for (int b0 = 0; b0 < rn; b0 += NA) {
                const int na = min(NA, rn - b0);
                const int abase = roff + b0;
                __syncthreads(); // previous batch's reads are done
                if (threadIdx.x < na) {
                    const int ang = abase + threadIdx.x;
                    const float st = gload(trig, 2 * ang), ct = gload(trig, 2 * ang + 1);
                    const float u00 = cx0 * ct - cy0 * st, u10 = cx1 * ct - cy0 * st;
                    const float u01 = cx0 * ct - cy1 * st, u11 = cx1 * ct - cy1 * st;
                    float umn = fminf(fminf(u00, u10), fminf(u01, u11)) + shift;
                    float umx = fmaxf(fmaxf(u00, u10), fmaxf(u01, u11)) + shift;
                    if (REF || clamp_u) {
                        umn = fminf(fmaxf(umn, edge0), edge1);
                        umx = fminf(fmaxf(umx, edge0), edge1);
                    }
                    const int umin = (int)floorf(umn - 0.5f) - 1;
                    if ((int)floorf(umx - 0.5f) + 2 - umin >= kBpWb)
                        atomicOr(err, 2);
                    s_umin[threadIdx.x] = umin;
                    s_trig[2 * threadIdx.x] = st;
                    s_trig[2 * threadIdx.x + 1] = ct;
                    if constexpr (LINE) {
                        const float *lt = uniform(S->line);
                        s_line[threadIdx.x] = make_float4(gload(lt, 8 * ang + 4), gload(lt, 8 * ang + 5),
                                                          gload(lt, 8 * ang + 6), gload(lt, 8 * ang + 7));
                    }
                    if constexpr (!REF) {
                        // u - 1/2 - umin at the super-tile's corner (joseph: less
                        // another 1/2, as its entries are found by rounding to
                        // nearest), rounded once; the line weights multiply
                        // coordinate errors by 1 / (M m)
                        s_base[threadIdx.x] = (float)((double)cx0 * (double)ct - (double)cy0 * (double)st +
                                                      (double)shift - (LINE ? 0.5 : 1.0) - (double)umin);
                    }
                }
                __syncthreads();
                {
                    // angle k = wave + 4 * (e / 2), bin i = lane + 64 * (e % 2);
                    // all loads first, then the LDS writes
                    constexpr int kPer = (NA / 4) * 2;
                    static_assert(kBpWb <= 128, "two bins per lane");
                    float v0[kPer], v1[kPer], y0[PAIR ? kPer : 1], y1[PAIR ? kPer : 1];
#pragma unroll
                    for (int e = 0; e < kPer; ++e) {
                        // clamped, always valid indices: no branches around the loads
                        const int k = min(wave + 4 * (e >> 1), na - 1);
                        const int i = lane + 64 * (e & 1);
                        const uint32_t rowo = (uint32_t)(abase + k) * (uint32_t)D;
                        const int bin = s_umin[k] + i;
                        if constexpr (REF || LINE) {
                            const int bc = min(max(bin, 0), D - 1);
                            const float t = gload(sino, rowo + (uint32_t)bc);
                            v0[e] = bin == bc ? t : 0.0f;
                            if constexpr (PAIR) {
                                const float t2 = gload(sino2, rowo + (uint32_t)bc);
                                y0[e] = bin == bc ? t2 : 0.0f;
                            }
                        } else {
                            // (t[i], t[i+1]) with clamp-to-edge bins; with clamp_u the
                            // coordinate is clamped instead, and the bins only to the row
                            const int lo = clamp_u ? 0 : bin_lo, hi = clamp_u ? D - 1 : bin_hi;
                            const int b0 = bin < lo ? lo : (bin > hi ? hi : bin);
                            const int b1 = bin + 1 < lo ? lo : (bin + 1 > hi ? hi : bin + 1);
                            v0[e] = gload(sino, rowo + (uint32_t)b0);
                            v1[e] = gload(sino, rowo + (uint32_t)b1);
                            if constexpr (PAIR) {
                                y0[e] = gload(sino2, rowo + (uint32_t)b0);
                                y1[e] = gload(sino2, rowo + (uint32_t)b1);
                            }
                        }
                    }
#pragma unroll
                    for (int e = 0; e < kPer; ++e) {
                        const int k = wave + 4 * (e >> 1);
                        const int i = lane + 64 * (e & 1);
                        if (k < na && i < kBpWb) {
                            if constexpr (LINE && PAIR) {
                                ldl[k * kBpWb + i] = make_float2(v0[e], y0[e]);
                            } else if constexpr (REF || LINE) {
                                lds[k * kBpWb + i] = v0[e];
                            } else {
                                // (e, d) with d = t[i+1] - t[i], e = t[i] - (i - 1/2) d: the
                                // value at window coordinate w in [i, i+1] is e + (w - 1/2) d,
                                // and rne_index(w - 1/2) is i or (at w = i or i + 1 exactly)
                                // a neighbour whose entry gives the same value
                                const float d = v1[e] - v0[e];
                                const float d2 = PAIR ? y1[e] - y0[e] : 0.0f;
                                const float ih = (float)i - 0.5f;
                                ldp[k * kBpWb + i] = make_entry<PAIR>(__builtin_fmaf(-ih, d, v0[e]), d,
                                                                      PAIR ? __builtin_fmaf(-ih, d2, y0[e]) : 0.0f, d2);
                            }
                        }
                    }
                }
                __syncthreads();
                for (int k = 0; k < na; ++k) {
                    const float st = s_trig[2 * k], ct = s_trig[2 * k + 1];
                    if constexpr (REF) {
                        const float *L = lds + k * kBpWb - s_umin[k];
                        const float fXc = fX * ct;
#pragma unroll
                        for (int g = 0; g < ROWS; g += 4) {
                            float t0[4], t1[4], a[4];
#pragma unroll
                            for (int h = 0; h < 4; ++h) {
                                const float fY = fY0 + (float)(4 * (g + h));
                                float u = fXc - fY * st + shift;
                                u = __builtin_amdgcn_fmed3f(u, edge0, edge1); // the clamp of :206-210
                                const float w = u - 0.5f;                     // >= 0: edge0 >= 0.5
                                const int i0 = cvt_flr(w);
                                a[h] = fract_pos(w);
                                t0[h] = L[i0];
                                t1[h] = L[i0 + 1];
                            }
#pragma unroll
                            for (int h = 0; h < 4; ++h) {
                                const float oma = 1.0f - a[h];
                                val[g + h] += oma * t0[h] + a[h] * t1[h];
                            }
                        }
                    } else {
                        // Window coordinates: row q's is fma(-ys_q, st, A), one FMA from
                        // the thread's column term A and the row's offset ys_q in the
                        // super-tile, whatever the tile's rows.  Joseph: less 1/2, so a
                        // sample's entry is rne_index of it, one float add
                        // (floor-converts and shifts issue at half the rate of float
                        // adds), and its address one shift-add.
                        const float A = __builtin_fmaf((float)lane, ct, s_base[k]);
                        if constexpr (LINE) {
                            const float4 L = s_line[k]; // (bp_k, bp_b0, bp_b1, 1/M)
                            if constexpr (PAIR) {
                                const float2 *B2 = ldl + k * kBpWb;
#pragma unroll
                                for (int g = 0; g < ROWS; g += 8) {
                                    float2 p[8], q[8];
                                    float fh[8];
#pragma unroll
                                    for (int h = 0; h < 8; ++h) {
                                        const float w = __builtin_fmaf(nys[g + h], st, A); // >= 1: inside the window
                                        const int i0 = cvt_flr(w);
                                        p[h] = B2[i0];
                                        q[h] = B2[i0 + 1];
                                        fh[h] = fract_pos(w) - 0.5f;
                                    }
#pragma unroll
                                    for (int h = 0; h < 8; ++h) {
                                        const float w0 =
                                            __builtin_amdgcn_fmed3f(__builtin_fmaf(fh[h], -L.x, L.y), 0.0f, L.w);
                                        const float w1 =
                                            __builtin_amdgcn_fmed3f(__builtin_fmaf(fh[h], L.x, L.z), 0.0f, L.w);
                                        acc[g + h] = __builtin_fmaf(w1, q[h].x, __builtin_fmaf(w0, p[h].x, acc[g + h]));
                                        acc2[g + h] =
                                            __builtin_fmaf(w1, q[h].y, __builtin_fmaf(w0, p[h].y, acc2[g + h]));
                                    }
                                }
                            } else {
                                const float *B = lds + k * kBpWb;
#pragma unroll
                                for (int g = 0; g < ROWS; g += 8) {
                                    float2 p[8];
                                    float fh[8];
#pragma unroll
                                    for (int h = 0; h < 8; ++h) {
                                        const float w = __builtin_fmaf(nys[g + h], st, A); // >= 1: inside the window
                                        const int i0 = cvt_flr(w);
                                        p[h] = make_float2(B[i0], B[i0 + 1]);
                                        fh[h] = fract_pos(w) - 0.5f;
                                    }
#pragma unroll
                                    for (int h = 0; h < 8; ++h) {
                                        const float w0 = __builtin_amdgcn_fmed3f(__builtin_fmaf(fh[h], -L.x, L.y), 0.0f, L.w);
                                        const float w1 = __builtin_amdgcn_fmed3f(__builtin_fmaf(fh[h], L.x, L.z), 0.0f, L.w);
                                        acc[g + h] = __builtin_fmaf(w1, p[h].y, __builtin_fmaf(w0, p[h].x, acc[g + h]));
                                    }
                                }
                            }
                        } else if (!clamp_u) {
                            const uint32_t P = lds_rne_base(ldp + k * kBpWb);
                            // a copy of A per branch: the compiler would otherwise hoist
                            // every row's coordinate above the branch, all live at once
                            float Ab = A;
                            asm volatile("" : "+v"(Ab));
#pragma unroll
                            for (int g = 0; g < ROWS; g += 8) {
                                E p[8];
                                float w[8];
#pragma unroll
                                for (int h = 0; h < 8; ++h) {
                                    w[h] = __builtin_fmaf(nys[g + h], st, Ab);
                                    p[h] = lds_rne_read<E>(P, w[h]);
                                }
#pragma unroll
                                for (int h = 0; h < 8; ++h) {
                                    acc[g + h] = __builtin_fmaf(w[h], p[h].y, acc[g + h] + p[h].x);
                                    if constexpr (PAIR)
                                        acc2[g + h] = __builtin_fmaf(w[h], p[h].w, acc2[g + h] + p[h].z);
                                }
                            }
                        } else {
                            // u clamped to [edge0, edge1] (:203-210), in window coordinates less 1/2
                            const float wlo = (edge0 - 1.0f) - (float)s_umin[k];
                            const float whi = (edge1 - 1.0f) - (float)s_umin[k];
                            const uint32_t P = lds_rne_base(ldp + k * kBpWb);
                            float Ab = A;
                            asm volatile("" : "+v"(Ab));
#pragma unroll
                            for (int g = 0; g < ROWS; g += 8) {
                                E p[8];
                                float w[8];
#pragma unroll
                                for (int h = 0; h < 8; ++h) {
                                    w[h] = __builtin_amdgcn_fmed3f(__builtin_fmaf(nys[g + h], st, Ab), wlo, whi);
                                    p[h] = lds_rne_read<E>(P, w[h]);
                                }
#pragma unroll
                                for (int h = 0; h < 8; ++h) {
                                    acc[g + h] = __builtin_fmaf(w[h], p[h].y, acc[g + h] + p[h].x);
                                    if constexpr (PAIR)
                                        acc2[g + h] = __builtin_fmaf(w[h], p[h].w, acc2[g + h] + p[h].z);
                                }
                            }
                        }
                    }
                }
            }
            if (REF) { // per-run accumulation (:216)
#pragma unroll
                for (int q = 0; q < ROWS; ++q) {
                    acc[q] = acc[q] + relax * val[q];
                    val[q] = 0.0f;
                }
            }
        }

        auto epilogue = [&](const uirk_slice *Sx, const float(&a)[ROWS]) {
            const TI *xin = uniform((const TI *)Sx->x);
            TI *xout = uniform((TI *)Sx->x_out);
            const float *pix_w = uniform(Sx->pix_w);
            const int posc = Sx->posc;
            const float gain = Sx->relax * Sx->pix_scale;
            const float *tv = REF ? nullptr : uniform(Sx->tv);
            const float tv_scale = Sx->tv_scale;
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
                const int py = item.y0 + wave + 4 * q;
                if (py >= N)
                    continue;
                const uint32_t o = (uint32_t)py * (uint32_t)N + (uint32_t)px;
                float out;
                if (REF) {
                    out = a[q];
                    if (pix_w) { // SIRT form (ufo-ir-sirt-task.c:208-212)
                        out = out * gload(pix_w, o);
                        out = out + gload(xin, o);
                    }
                } else if (LINE && tv) {
                    // split-Bregman sweep of the line model: P = 1 / (mu BP(1) + 8 lambda)
                    // per pixel, from C = 1 / BP(1) (0: BP(1) below the floor)
                    const float C = gload(pix_w, o), mu = Sx->pix_scale, lam = Sx->tv_scale;
                    const float P = C > 0.0f ? C / __builtin_fmaf(8.0f * lam, C, mu) : 1.0f / (8.0f * lam);
                    out = __builtin_fmaf(Sx->relax * P, __builtin_fmaf(mu, a[q], lam * gload(tv, o)),
                                         gload(xin, o));
                } else if (LINE && pix_w) {
                    out = __builtin_fmaf(a[q] * gain, gload(pix_w, o), gload(xin, o));
                } else {
                    out = __builtin_fmaf(a[q], gain, gload(xin, o));
                    if (tv) // split-Bregman TV term, evaluated at the same x
                        out = __builtin_fmaf(tv_scale, gload(tv, o), out);
                }
                if (posc)
                    out = out > 0 ? out : 0.0f;
                gstore(xout, o, out);
            }
        };
        if (px < N) {
            epilogue(S, acc);
            if constexpr (PAIR) {
                if (paired)
                    epilogue(S2, acc2);
            }
        }
    }
}

// ------------------------------------------------------------------------
// Fused SART step: the BP of angle k - 1 and the FP of angle k in one pass
// ------------------------------------------------------------------------
// Block (t, s) owns the 64 x 64 tile t of slice s.  It stages the tile with a
// one-pixel halo (66 x 66) from x_in, applies the BP of angle bp_angle to
// every staged pixel (x += relax * A^T r[bp_angle], fast-path arithmetic:
// exact per-pixel detector coordinate, clamp-to-edge bins as k_bp_tab),
// writes the 64 x 64 interior to x_out (a different buffer: the
// halo belongs to the neighbours) and keeps the updated tile in LDS.  It then
// sums, for angle fp_angle, every Joseph sample whose left (lower) tap c0
// lies in the tile's positions (tile 0 also owns c0 = -1) over the tile's 64
// lines, ray by ray.  Each ray's partial goes to part[band][tile % 3][ray]:
// a ray's samples in one band of lines touch at most three consecutive tiles
// (even with the margins below), so tiles sharing a slot never share a ray.
// k_sart_rows then adds the partials of the tiles each ray crosses.
//
// The sample-to-tile assignment must not depend on the tile: every tile
// evaluates the tap position of (ray, line) as fma(line, step, q0) in float,
// q0 the ray's position at line 0 rounded once from double, and visits only
// the lines whose taps can fall in its positions (with a one-position margin).
constexpr int kSt = 64;          // tile edge
constexpr int kSl = kSt + 2;     // staged edge
constexpr int kSartRays = 128;   // candidate rays per tile (at most 98 by construction)

// Candidate rays of a tile: those whose tap position c = pos - 0.5 falls in
// [p_base - 3, p_base + 67) on a line of [la, lb]; pos(s, l) = (s + 0.5 - axis)
// det_step + C + l step is affine in (s, l), so the set is the interval between
// the corner solutions (one more ray of margin each side).
__device__ __forceinline__ void sart_candidates(double det_step, double step, double C, double axis, int la, int lb,
                                                int p_base, int D, int &s_lo, int &s_hi)
{
    double smin = 1e300, smax = -1e300;
    const int ls[2] = {la, lb};
    const double cs[2] = {(double)p_base - 3.0, (double)p_base + 67.0};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const double s = (cs[j] + 0.5 - C - (double)ls[i] * step) / det_step - 0.5 + axis;
            smin = fmin(smin, s);
            smax = fmax(smax, s);
        }
    s_lo = max((int)floor(smin) - 1, 0);
    s_hi = min((int)ceil(smax) + 1, D - 1);
}

template <typename T, typename TI>
__global__ __launch_bounds__(256) void k_sart_step(const uirk_sart_slice *__restrict__ slices,
                                                   const uirk_asd_state *__restrict__ state, int bp_angle,
                                                   int fp_angle, int *__restrict__ err)
{
    __shared__ float tile[kSl * kSl];
    __shared__ float half_sum[kSartRays];
    const uirk_sart_slice *S = &slices[blockIdx.y];
    const int N = S->n, D = S->n_dets;
    const int nt = (N + kSt - 1) / kSt;
    if ((int)blockIdx.x >= nt * nt)
        return; // uniform over the block
    const int tx = (int)blockIdx.x % nt, ty = (int)blockIdx.x / nt;
    const int x0 = tx * kSt, y0 = ty * kSt;
    const TI *xin = uniform((const TI *)S->x_in);
    TI *xout = uniform((TI *)S->x_out);
    const float *trig = uniform(S->trig);
    const float axis_f = S->axis;
    const float required_width = axis_f * 2;
    const float diff = (float)D - required_width;
    const float origin_shift = (float)N / 2.0f;

    // BP of bp_angle (projector-parallel-joseph.cl:174-190 geometry, clamp-to-edge bins)
    const bool do_bp = bp_angle >= 0 && bp_angle < S->n_angles;
    float bst = 0.0f, bct = 0.0f, gain = 0.0f;
    float w_lo = 0.0f, w_hi = 0.0f;
    const T *row = nullptr;
    if (do_bp) {
        bst = gload(trig, 2 * bp_angle);
        bct = gload(trig, 2 * bp_angle + 1);
        const float half_active = diff < 0 ? (float)D - axis_f : axis_f;
        // the clamp window of u - 0.5 (:180-182, :203-210)
        w_lo = (axis_f - half_active + 0.5f) - 0.5f;
        w_hi = (axis_f + half_active - 0.5f) - 0.5f;
        gain = state[blockIdx.y].relax;
        row = uniform((const T *)S->r) + (size_t)bp_angle * D;
    }
    const float shift = (float)D / 2.0f - diff / 2.0f;
    // The BP taps of one pixel: u = X cos - Y sin + axis - 0.5, clamped to the
    // active window; the bins are also kept inside the row whatever the axis.
    auto bp_taps = [&](int px, int py, float &a, float &t0, float &t1) {
        const float fX = (float)px + 0.5f - origin_shift, fY = (float)py + 0.5f - origin_shift;
        const float u = __builtin_amdgcn_fmed3f(__builtin_fmaf(-fY, bst, __builtin_fmaf(fX, bct, shift - 0.5f)),
                                                w_lo, w_hi);
        const int bin = cvt_flr(u);
        a = u - floorf(u);
        const int b0 = min(max(bin, 0), D - 1), b1 = min(max(bin + 1, 0), D - 1);
        t0 = gload(row, (uint32_t)b0);
        t1 = gload(row, (uint32_t)b1);
    };
    // Staged (i, j) is pixel (y0 - 1 + i, x0 - 1 + j); out-of-image entries are 0.
    // Interior: one column and 16 rows per thread, in two batches of 8 rows
    // whose loads are all issued first.
    {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int px = x0 + lane;
        const bool col_in = px < N;
        constexpr int kRows = 8;
        for (int k0 = 0; k0 < kSt / 4; k0 += kRows) {
            float xv[kRows], t0[kRows], t1[kRows], a[kRows];
#pragma unroll
            for (int k = 0; k < kRows; ++k) {
                const int py = min(y0 + wave + 4 * (k0 + k), N - 1), pxc = min(px, N - 1);
                xv[k] = gload(xin, (uint32_t)py * (uint32_t)N + (uint32_t)pxc);
                a[k] = t0[k] = t1[k] = 0.0f;
                if (do_bp)
                    bp_taps(pxc, py, a[k], t0[k], t1[k]);
            }
#pragma unroll
            for (int k = 0; k < kRows; ++k) {
                const int py = y0 + wave + 4 * (k0 + k);
                float v = 0.0f;
                if (col_in && py < N) {
                    v = do_bp ? __builtin_fmaf(__builtin_fmaf(a[k], t1[k] - t0[k], t0[k]), gain, xv[k]) : xv[k];
                    gstore(xout, (uint32_t)py * (uint32_t)N + (uint32_t)px, v);
                }
                tile[(wave + 4 * (k0 + k) + 1) * kSl + lane + 1] = v;
            }
        }
    }
    // Halo ring (4 * 65 entries): updated here, written by the neighbours.
    for (int e = threadIdx.x; e < 4 * (kSl - 1); e += 256) {
        int i, j;
        if (e < kSl) {
            i = 0;
            j = e;
        } else if (e < 2 * kSl) {
            i = kSl - 1;
            j = e - kSl;
        } else if (e < 2 * kSl + kSt) {
            i = e - 2 * kSl + 1;
            j = 0;
        } else {
            i = e - 2 * kSl - kSt + 1;
            j = kSl - 1;
        }
        const int py = y0 - 1 + i, px = x0 - 1 + j;
        float v = 0.0f;
        if (py >= 0 && py < N && px >= 0 && px < N) {
            v = gload(xin, (uint32_t)py * (uint32_t)N + (uint32_t)px);
            if (do_bp) {
                float a, t0, t1;
                bp_taps(px, py, a, t0, t1);
                v = __builtin_fmaf(__builtin_fmaf(a, t1 - t0, t0), gain, v);
            }
        }
        tile[i * kSl + j] = v;
    }
    if (fp_angle < 0 || fp_angle >= S->n_angles)
        return; // uniform over the block
    __syncthreads();

    // FP of fp_angle over the tile (the fast FP's geometry, exact anchors)
    const float st = gload(trig, 2 * fp_angle), ct = gload(trig, 2 * fp_angle + 1);
    const bool hor = !(fabsf(st) <= fabsf(ct));
    const double det_step = hor ? -1.0 / (double)st : 1.0 / (double)ct;
    const double step = hor ? (double)ct / (double)st : (double)st / (double)ct;
    const double C = -0.5 * (double)N * step + 0.5 * (double)N;
    const int l0 = hor ? x0 : y0, p_base = hor ? y0 : x0;
    const int nl = min(kSt, N - l0);
    int s_lo, s_hi;
    sart_candidates(det_step, step, C, (double)axis_f, l0, l0 + nl - 1, p_base, D, s_lo, s_hi);
    if (s_hi - s_lo + 1 > kSartRays && threadIdx.x == 0)
        atomicOr(err, 8); // cannot happen (at most 98): reported, never silent
    const int r = threadIdx.x & (kSartRays - 1), h = threadIdx.x / kSartRays;
    const int s = s_lo + r;
    const float rot = diff / 2.0f;
    const bool cand = s <= s_hi;
    const bool on = cand && !((diff < 0 && (float)s < fabsf(diff)) || (diff > 0 && (float)s > required_width));
    float acc = 0.0f;
    if (on) {
        const double tdet = ((0.5 + (double)rot) + (double)s) - (double)(0.5f * (float)D);
        const double q0 = tdet * det_step + C - 0.5; // tap coordinate at line 0
        const float q0f = (float)q0, se = (float)step;
        const int c_min = p_base == 0 ? -1 : p_base, c_max = p_base + kSt - 1;
        int lb = l0 + 32 * h, le = min(lb + 32, l0 + nl);
        if (fabs(step) > 1e-9) { // lines whose tap is within one position of [c_min, c_max + 1)
            const double la = ((double)c_min - 1.0 - q0) / step, lz = ((double)c_max + 2.0 - q0) / step;
            lb = max(lb, (int)floor(fmin(la, lz)) - 1);
            le = min(le, (int)ceil(fmax(la, lz)) + 2);
        }
        for (int l = lb; l < le; ++l) {
            const float qa = __builtin_fmaf((float)l, se, q0f);
            const int c0 = cvt_flr(qa);
            if (c0 >= c_min && c0 <= c_max) {
                const float a = __builtin_amdgcn_fractf(qa);
                const int i0 = c0 - p_base + 1, li = l - l0 + 1; // staged indices, i0 in [0, 64]
                const float t0 = hor ? tile[i0 * kSl + li] : tile[li * kSl + i0];
                const float t1 = hor ? tile[(i0 + 1) * kSl + li] : tile[li * kSl + i0 + 1];
                acc = __builtin_fmaf(a, t1 - t0, acc + t0);
            }
        }
    }
    if (h == 1)
        half_sum[r] = acc;
    __syncthreads();
    if (h == 0 && cand) {
        const int band = hor ? tx : ty, slot = (hor ? ty : tx) % 3;
        S->part[((size_t)band * 3 + (size_t)slot) * (size_t)D + (size_t)s] = acc + half_sum[r];
    }
}

// Residual rows of the fused step: r[angle] = (b[angle] + corr * FP) * R over
// every slice that has the angle; a ray's FP is the sum, band by band in
// order, of the partials of the tiles its taps cross (with one position of
// margin; every such tile had the ray among its candidates).
template <typename T>
__global__ __launch_bounds__(256) void k_sart_rows(const uirk_sart_slice *__restrict__ slices, int angle)
{
    const uirk_sart_slice *S = &slices[blockIdx.y];
    const int N = S->n, D = S->n_dets;
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= D || angle >= S->n_angles)
        return;
    const float *trig = uniform(S->trig);
    const float st = gload(trig, 2 * angle), ct = gload(trig, 2 * angle + 1);
    const bool hor = !(fabsf(st) <= fabsf(ct));
    const float axis_f = S->axis;
    const float required_width = axis_f * 2;
    const float diff = (float)D - required_width;
    const float rot = diff / 2.0f;
    const bool on = !((diff < 0 && (float)s < fabsf(diff)) || (diff > 0 && (float)s > required_width));
    const uint32_t o = (uint32_t)angle * (uint32_t)D + (uint32_t)s;
    const float bval = gload(uniform((const T *)S->b), o);
    float w = bval;
    if (on) {
        const double det_step = hor ? -1.0 / (double)st : 1.0 / (double)ct;
        const double step = hor ? (double)ct / (double)st : (double)st / (double)ct;
        const double C = -0.5 * (double)N * step + 0.5 * (double)N;
        const double tdet = ((0.5 + (double)rot) + (double)s) - (double)(0.5f * (float)D);
        const double p0 = tdet * det_step + C;
        const int nt = (N + kSt - 1) / kSt;
        const float *part = uniform(S->part);
        float sum = 0.0f;
        for (int band = 0; band < nt; ++band) {
            const int la = band * kSt, lb = min(la + kSt - 1, N - 1);
            const double ca = p0 + (double)la * step - 0.5, cb = p0 + (double)lb * step - 0.5;
            int lo = (int)floor(fmin(ca, cb)) - 1, hi = (int)floor(fmax(ca, cb)) + 1;
            if (hi < -1 || lo > N - 1)
                continue;
            lo = max(lo, -1);
            hi = min(hi, N - 1);
            const int t_lo = lo < 0 ? 0 : lo / kSt, t_hi = hi < 0 ? 0 : hi / kSt;
            for (int t = t_lo; t <= t_hi; ++t)
                sum += part[((size_t)band * 3 + (size_t)(t % 3)) * (size_t)D + (size_t)s];
        }
        w = bval + sum * S->corr;
    }
    const float *ray_w = uniform(S->ray_w);
    if (ray_w)
        w = w * gload(ray_w, o);
    gstore(uniform((T *)S->r), o, w);
}

template <typename T, typename TI>
int launch_sart(const uirk_sart_slice *sl, const uirk_asd_state *state, int n_slices, int max_n, int bp_angle,
                int fp_angle, int *err, hipStream_t st)
{
    const int t = (max_n + kSt - 1) / kSt;
    hipLaunchKernelGGL((k_sart_step<T, TI>), dim3(t * t, n_slices), dim3(256), 0, st, sl, state, bp_angle, fp_angle,
                       err);
    return (int)hipGetLastError();
}

__global__ void k_convert_f32_to(const float *src, void *dst, int dtype, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (dtype == UIR_DTYPE_F16)
            ((__half *)dst)[i] = __float2half(src[i]);
        else if (dtype == UIR_DTYPE_BF16)
            ((__hip_bfloat16 *)dst)[i] = __float2bfloat16(src[i]);
        else
            ((float *)dst)[i] = src[i];
    }
}

__global__ void k_convert_to_f32(const void *src, int dtype, float *dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (dtype == UIR_DTYPE_F16)
            dst[i] = __half2float(((const __half *)src)[i]);
        else if (dtype == UIR_DTYPE_BF16)
            dst[i] = __bfloat162float(((const __hip_bfloat16 *)src)[i]);
        else
            dst[i] = ((const float *)src)[i];
    }
}

// Split-Bregman TV term of the u-subproblem, tv = Dx^T (ex - Dx u) + Dy^T (ey - Dy u),
// with the periodic differences of ufo-ir-gradient-processor.cl:1-51
// ((Dx u)[x] = u[x] - u[x-1], (Dx^T v)[x] = v[x] - v[x+1]).  One thread per
// pixel, rows of 256 pixels per block; neighbours come through L2.
template <typename T>
__global__ __launch_bounds__(256) void k_sb_tv(const T *__restrict__ u, const float *__restrict__ ex,
                                               const float *__restrict__ ey, float *__restrict__ tv, uint32_t n)
{
    const uint32_t x = blockIdx.x * 256u + threadIdx.x, y = blockIdx.y;
    if (x >= n)
        return;
    const uint32_t xl = x == 0 ? n - 1 : x - 1, xr = x + 1 == n ? 0 : x + 1;
    const uint32_t yu = y == 0 ? n - 1 : y - 1, yd = y + 1 == n ? 0 : y + 1;
    const uint32_t o = y * n + x;
    const float c = gload(u, o);
    // h(p) = e(p) - (u(p) - u(p - 1)) along each axis; tv = h(p) - h(p + 1)
    const float hx0 = ex[o] - (c - gload(u, y * n + xl));
    const float hx1 = ex[y * n + xr] - (gload(u, y * n + xr) - c);
    const float hy0 = ey[o] - (c - gload(u, yu * n + x));
    const float hy1 = ey[yd * n + x] - (gload(u, yd * n + x) - c);
    tv[o] = (hx0 - hx1) + (hy0 - hy1);
}

// Split-Bregman shrinkage and Bregman update (ufo-ir-sbtv-task.c:360-402 form):
// g = D u + b, t = max(|g| - 1/lambda, 0) / max(|g|, 1e-12), d = t g,
// b <- g - d; stores b and e = d - b.
template <typename T>
__global__ __launch_bounds__(256) void k_sb_shrink(const T *__restrict__ u, float *__restrict__ bx,
                                                   float *__restrict__ by, float *__restrict__ ex,
                                                   float *__restrict__ ey, uint32_t n, float inv_lambda)
{
    const uint32_t x = blockIdx.x * 256u + threadIdx.x, y = blockIdx.y;
    if (x >= n)
        return;
    const uint32_t xl = x == 0 ? n - 1 : x - 1, yu = y == 0 ? n - 1 : y - 1;
    const uint32_t o = y * n + x;
    const float c = gload(u, o);
    const float gx = (c - gload(u, y * n + xl)) + bx[o];
    const float gy = (c - gload(u, yu * n + x)) + by[o];
    const float s = sqrtf(gx * gx + gy * gy);
    const float t = fmaxf(s - inv_lambda, 0.0f) / fmaxf(s, 1e-12f);
    const float dx = t * gx, dy = t * gy;
    bx[o] = gx - dx;
    by[o] = gy - dy;
    ex[o] = dx - (gx - dx);
    ey[o] = dy - (gy - dy);
}

// Ray weights R = 1 / FP(1) of the fast path's geometry, FP(1) in FP64 with
// exact coordinates: one thread per ray, every line, the operations of the
// FP64 evaluation of the projector (samples outside the image add exact
// zeros).  R = 0 where FP(1) = 0 or FP(1) < r_min.  FP64 keeps the floor's
// decision from depending on FP32 rounding: rays that clip a corner have
// FP(1) made of one or two fractional weights, which float coordinates
// near 1000 shift by ~1e-4.
__global__ __launch_bounds__(256) void k_ray_weights64(const float *__restrict__ trig, int N, int D, int A,
                                                       float axis_f, float r_min, float *__restrict__ ray_w)
{
    const int s = blockIdx.x * 256 + threadIdx.x, ang = blockIdx.y;
    if (s >= D || ang >= A)
        return;
    const double axis = (double)axis_f;
    const double required_width = 2.0 * axis;
    const double diff = (double)D - required_width;
    float w = 0.0f;
    if (!((diff < 0 && (double)s < fabs(diff)) || (diff > 0 && (double)s > required_width))) {
        const float stf = trig[2 * ang], ctf = trig[2 * ang + 1];
        const double st = (double)stf, ct = (double)ctf;
        const bool vertical = fabsf(stf) <= fabsf(ctf);
        const double t = (double)s + 0.5 - axis;
        const double det_step = vertical ? 1.0 / ct : -1.0 / st;
        const double step = vertical ? st / ct : ct / st;
        const double v0 = t * det_step - 0.5 * N * step + 0.5 * N;
        double acc = 0.0;
        for (int i = 0; i < N; ++i) {
            const double v = v0 + i * step - 0.5;
            const double fv = floor(v);
            const int j0 = (int)fv;
            const double b = v - fv;
            const double t0 = (j0 >= 0 && j0 < N) ? 1.0 : 0.0, t1 = (j0 + 1 >= 0 && j0 + 1 < N) ? 1.0 : 0.0;
            acc += (1.0 - b) * t0 + b * t1;
        }
        w = (acc != 0.0 && acc >= (double)r_min) ? (float)(1.0 / acc) : 0.0f;
    }
    ray_w[(size_t)ang * D + s] = w;
}

template <typename T, typename TI, bool REF, int APT>
int launch_fp(const uirk_slice *sl, const uirk_slice &one, const uirk_fp_item *items, int n,
              const uirk_xcd_part &part, int grid, int *err, hipStream_t st)
{
    hipLaunchKernelGGL((k_fp_tab<T, TI, REF, APT>), dim3(grid), dim3(256), 0, st, sl, one, items, n, part, err);
    return (int)hipGetLastError();
}

template <typename T, typename TI, int APT, bool LINE = false>
int launch_fp_pairs(const uirk_slice *sl, const uirk_slice &one, const uirk_fp_item *items, int n,
                    const uirk_xcd_part &part, int grid, int *err, hipStream_t st)
{
    hipLaunchKernelGGL((k_fp_pairs<T, TI, APT, LINE>), dim3(grid), dim3(256), 0, st, sl, one, items, n, part, err);
    return (int)hipGetLastError();
}

template <typename T, typename TI, bool REF, int NA, int ROWS, bool LINE = false, bool PAIR = false>
int launch_bp(const uirk_slice *sl, const uirk_slice &one, const uirk_bp_item *items, int n,
              const uirk_xcd_part &part, int grid, int *err, hipStream_t st)
{
    hipLaunchKernelGGL((k_bp_tab<T, TI, REF, NA, ROWS, LINE, PAIR>), dim3(grid), dim3(256), 0, st, sl, one, items, n,
                       part, err);
    return (int)hipGetLastError();
}

// fast-path FP over the tile variants: 16, 8 or 32 angles per item (4, 2 or 8 per thread);
// the line model takes the default tile
template <typename T, typename TI>
int launch_fp_fast(const uirk_tile &tile, bool line, bool, const uirk_slice *sl, const uirk_slice &one,
                   const uirk_fp_item *items, int n, const uirk_xcd_part &part, int grid, int *err, hipStream_t st)
{
    if (line)
        return launch_fp_pairs<T, TI, 4, true>(sl, one, items, n, part, grid, err, st);
    if (tile.fp_single) {
        if (tile.fp_angles == 8)
            return launch_fp<T, TI, false, 2>(sl, one, items, n, part, grid, err, st);
        if (tile.fp_angles == 32)
            return launch_fp<T, TI, false, 8>(sl, one, items, n, part, grid, err, st);
        return launch_fp<T, TI, false, 4>(sl, one, items, n, part, grid, err, st);
    }
    if (tile.fp_angles == 8)
        return launch_fp_pairs<T, TI, 2>(sl, one, items, n, part, grid, err, st);
    if (tile.fp_angles == 32)
        return launch_fp_pairs<T, TI, 8>(sl, one, items, n, part, grid, err, st);
    return launch_fp_pairs<T, TI, 4>(sl, one, items, n, part, grid, err, st);
}

// fast-path BP over the tile variants: 32 or 16 staged angles, tiles of 64 or 32 rows
template <typename T, typename TI>
int launch_bp_fast(const uirk_tile &tile, bool line, bool pair, const uirk_slice *sl, const uirk_slice &one,
                   const uirk_bp_item *items, int n, const uirk_xcd_part &part, int grid, int *err, hipStream_t st)
{
    if (line && pair) // 32 staged angles; tiles of 64 or 32 rows
        return tile.bp_rows == 32 ? launch_bp<T, TI, false, 32, 8, true, true>(sl, one, items, n, part, grid, err, st)
                                  : launch_bp<T, TI, false, 32, 16, true, true>(sl, one, items, n, part, grid, err, st);
    if (line)
        return tile.bp_rows == 32 ? launch_bp<T, TI, false, 32, 8, true>(sl, one, items, n, part, grid, err, st)
                                  : launch_bp<T, TI, false, 32, 16, true>(sl, one, items, n, part, grid, err, st);
    const bool na16 = tile.bp_angles == 16, r8 = tile.bp_rows == 32;
    if (pair) {
        if (na16 && r8)
            return launch_bp<T, TI, false, 16, 8, false, true>(sl, one, items, n, part, grid, err, st);
        if (na16)
            return launch_bp<T, TI, false, 16, 16, false, true>(sl, one, items, n, part, grid, err, st);
        if (r8)
            return launch_bp<T, TI, false, 32, 8, false, true>(sl, one, items, n, part, grid, err, st);
        return launch_bp<T, TI, false, 32, 16, false, true>(sl, one, items, n, part, grid, err, st);
    }
    if (na16 && r8)
        return launch_bp<T, TI, false, 16, 8>(sl, one, items, n, part, grid, err, st);
    if (na16)
        return launch_bp<T, TI, false, 16, 16>(sl, one, items, n, part, grid, err, st);
    if (r8)
        return launch_bp<T, TI, false, 32, 8>(sl, one, items, n, part, grid, err, st);
    return launch_bp<T, TI, false, 32, 16>(sl, one, items, n, part, grid, err, st);
}

// (sinogram, image) storage pairs: both FP32, both FP16, both BF16, or
// FP16 / BF16 sinograms with an FP32 image
template <bool FP>
int launch_fast(int dtype, int img_dtype, const uirk_tile &tile, bool line, bool pair, const uirk_slice *sl,
                const uirk_slice &one, const void *items, int n, const uirk_xcd_part &part, int grid, int *err,
                hipStream_t st)
{
#define UIRK_LAUNCH(TS, TIM)                                                                                         \
    return FP ? launch_fp_fast<TS, TIM>(tile, line, pair, sl, one, (const uirk_fp_item *)items, n, part, grid, err, \
                                        st)                                                                          \
              : launch_bp_fast<TS, TIM>(tile, line, pair, sl, one, (const uirk_bp_item *)items, n, part, grid, err, \
                                        st)
    if (img_dtype == UIR_DTYPE_F32) {
        switch (dtype) {
        case UIR_DTYPE_F16: UIRK_LAUNCH(__half, float);
        case UIR_DTYPE_BF16: UIRK_LAUNCH(__hip_bfloat16, float);
        case UIR_DTYPE_F32: UIRK_LAUNCH(float, float);
        default: return (int)hipErrorInvalidValue;
        }
    }
    if (img_dtype != dtype)
        return (int)hipErrorInvalidValue;
    switch (dtype) {
    case UIR_DTYPE_F16: UIRK_LAUNCH(__half, __half);
    case UIR_DTYPE_BF16: UIRK_LAUNCH(__hip_bfloat16, __hip_bfloat16);
    default: return (int)hipErrorInvalidValue;
    }
#undef UIRK_LAUNCH
}

bool tile_ok(const uirk_tile *t)
{
    return (t->fp_angles == 16 || t->fp_angles == 8 || t->fp_angles == 32) && (t->bp_angles == 32 || t->bp_angles == 16) &&
           (t->bp_rows == 64 || t->bp_rows == 32) && (t->fp_single == 0 || t->fp_single == 1);
}

} // namespace

extern "C" {

int uirk_fp_table(const uirk_slice *d_slices, const uirk_slice *one, const uirk_fp_item *d_items, int n_items,
                  const uirk_xcd_part *part, int mode, int dtype, int img_dtype, const uirk_tile *tile, int grid,
                  int *d_err, void *stream)
{
    uirk_slice single{};
    if (!d_slices) {
        if (!one)
            return (int)hipErrorInvalidValue;
        single = *one;
    }
    const uirk_tile tl = tile ? *tile : uirk_tile UIRK_TILE_DEFAULT;
    if (!tile_ok(&tl))
        return (int)hipErrorInvalidValue;
    if (n_items <= 0)
        return 0;
    if (grid <= 0 || grid > n_items)
        grid = n_items;
    uirk_xcd_part pt{};
    if (part)
        pt = *part;
    hipStream_t st = (hipStream_t)stream;
    if (mode == 0) {
        if (dtype != UIR_DTYPE_F32 || img_dtype != UIR_DTYPE_F32 || tl.fp_angles != UIRK_FP_ANGLES)
            return (int)hipErrorInvalidValue;
        return launch_fp<float, float, true, 4>(d_slices, single, d_items, n_items, pt, grid, d_err, st);
    }
    if (mode != 1 && mode != 2)
        return (int)hipErrorInvalidValue;
    return launch_fast<true>(dtype, img_dtype, tl, mode == 2, false, d_slices, single, d_items, n_items, pt, grid,
                             d_err, st);
}

int uirk_ray_weights64(const float *d_trig, int n, int n_dets, int n_angles, float axis, float r_min, float *d_ray_w,
                       void *stream)
{
    if (n <= 0 || n_dets <= 0 || n_angles <= 0)
        return 0;
    if (n_angles > 65535)
        return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ray_weights64, dim3((n_dets + 255) / 256, n_angles), dim3(256), 0, (hipStream_t)stream,
                       d_trig, n, n_dets, n_angles, axis, r_min, d_ray_w);
    return (int)hipGetLastError();
}

int uirk_sart_step(const uirk_sart_slice *d_slices, const uirk_asd_state *d_state, int n_slices, int max_n,
                   int dtype, int img_dtype, int bp_angle, int fp_angle, int *d_err, void *stream)
{
    if (!d_slices || !d_state || n_slices > 65535 || max_n > 16384)
        return (int)hipErrorInvalidValue;
    if (n_slices <= 0 || max_n <= 0)
        return 0;
    hipStream_t st = (hipStream_t)stream;
    if (img_dtype == UIR_DTYPE_F32) {
        switch (dtype) {
        case UIR_DTYPE_F16:
            return launch_sart<__half, float>(d_slices, d_state, n_slices, max_n, bp_angle, fp_angle, d_err, st);
        case UIR_DTYPE_BF16:
            return launch_sart<__hip_bfloat16, float>(d_slices, d_state, n_slices, max_n, bp_angle, fp_angle, d_err,
                                                      st);
        case UIR_DTYPE_F32:
            return launch_sart<float, float>(d_slices, d_state, n_slices, max_n, bp_angle, fp_angle, d_err, st);
        default: return (int)hipErrorInvalidValue;
        }
    }
    if (img_dtype != dtype)
        return (int)hipErrorInvalidValue;
    switch (dtype) {
    case UIR_DTYPE_F16:
        return launch_sart<__half, __half>(d_slices, d_state, n_slices, max_n, bp_angle, fp_angle, d_err, st);
    case UIR_DTYPE_BF16:
        return launch_sart<__hip_bfloat16, __hip_bfloat16>(d_slices, d_state, n_slices, max_n, bp_angle, fp_angle,
                                                           d_err, st);
    default: return (int)hipErrorInvalidValue;
    }
}

int uirk_sart_rows(const uirk_sart_slice *d_slices, int n_slices, int max_d, int dtype, int angle, void *stream)
{
    if (!d_slices || angle < 0 || n_slices > 65535)
        return (int)hipErrorInvalidValue;
    if (n_slices <= 0 || max_d <= 0)
        return 0;
    const dim3 grid((max_d + 255) / 256, n_slices);
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
    case UIR_DTYPE_F16: hipLaunchKernelGGL(k_sart_rows<__half>, grid, dim3(256), 0, st, d_slices, angle); break;
    case UIR_DTYPE_BF16:
        hipLaunchKernelGGL(k_sart_rows<__hip_bfloat16>, grid, dim3(256), 0, st, d_slices, angle);
        break;
    case UIR_DTYPE_F32: hipLaunchKernelGGL(k_sart_rows<float>, grid, dim3(256), 0, st, d_slices, angle); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

int uirk_bp_table(const uirk_slice *d_slices, const uirk_slice *one, const uirk_bp_item *d_items, int n_items,
                  const uirk_xcd_part *part, int mode, int dtype, int img_dtype, const uirk_tile *tile, int grid,
                  int *d_err, void *stream)
{
    uirk_slice single{};
    if (!d_slices) {
        if (!one)
            return (int)hipErrorInvalidValue;
        single = *one;
    }
    const uirk_tile tl = tile ? *tile : uirk_tile UIRK_TILE_DEFAULT;
    if (!tile_ok(&tl))
        return (int)hipErrorInvalidValue;
    if (n_items <= 0)
        return 0;
    if (grid <= 0 || grid > n_items)
        grid = n_items;
    uirk_xcd_part pt{};
    if (part)
        pt = *part;
    hipStream_t st = (hipStream_t)stream;
    if (mode == 0) {
        if (dtype != UIR_DTYPE_F32 || img_dtype != UIR_DTYPE_F32 || tl.bp_angles != 32 || tl.bp_rows != UIRK_BP_TILE)
            return (int)hipErrorInvalidValue;
        return launch_bp<float, float, true, 32, 16>(d_slices, single, d_items, n_items, pt, grid, d_err, st);
    }
    if (mode < 1 || mode > 4 || (mode >= 3 && !d_slices))
        return (int)hipErrorInvalidValue;
    return launch_fast<false>(dtype, img_dtype, tl, mode == 2 || mode == 4, mode >= 3, d_slices, single, d_items,
                              n_items, pt, grid, d_err, st);
}

int uirk_sb_tv(const void *u, int dtype, const float *ex, const float *ey, float *tv, uint32_t n, void *stream)
{
    if (!n)
        return 0;
    const dim3 grid((n + 255) / 256, n);
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
    case UIR_DTYPE_F16:
        hipLaunchKernelGGL(k_sb_tv<__half>, grid, dim3(256), 0, st, (const __half *)u, ex, ey, tv, n);
        break;
    case UIR_DTYPE_BF16:
        hipLaunchKernelGGL(k_sb_tv<__hip_bfloat16>, grid, dim3(256), 0, st, (const __hip_bfloat16 *)u, ex, ey, tv, n);
        break;
    default: hipLaunchKernelGGL(k_sb_tv<float>, grid, dim3(256), 0, st, (const float *)u, ex, ey, tv, n);
    }
    return (int)hipGetLastError();
}

int uirk_sb_shrink(const void *u, int dtype, float *bx, float *by, float *ex, float *ey, uint32_t n, float lambda,
                   void *stream)
{
    if (!n)
        return 0;
    const dim3 grid((n + 255) / 256, n);
    hipStream_t st = (hipStream_t)stream;
    const float il = 1.0f / lambda;
    switch (dtype) {
    case UIR_DTYPE_F16:
        hipLaunchKernelGGL(k_sb_shrink<__half>, grid, dim3(256), 0, st, (const __half *)u, bx, by, ex, ey, n, il);
        break;
    case UIR_DTYPE_BF16:
        hipLaunchKernelGGL(k_sb_shrink<__hip_bfloat16>, grid, dim3(256), 0, st, (const __hip_bfloat16 *)u, bx, by, ex,
                           ey, n, il);
        break;
    default: hipLaunchKernelGGL(k_sb_shrink<float>, grid, dim3(256), 0, st, (const float *)u, bx, by, ex, ey, n, il);
    }
    return (int)hipGetLastError();
}

int uirk_convert(const void *src, int src_dtype, void *dst, int dst_dtype, size_t n, void *stream)
{
    if (!n)
        return 0;
    size_t g = (n + 255) / 256;
    if (g > 65536)
        g = 65536;
    if (src_dtype == UIR_DTYPE_F32)
        hipLaunchKernelGGL(k_convert_f32_to, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream,
                           (const float *)src, dst, dst_dtype, n);
    else if (dst_dtype == UIR_DTYPE_F32)
        hipLaunchKernelGGL(k_convert_to_f32, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, src, src_dtype,
                           (float *)dst, n);
    else
        return (int)hipErrorInvalidValue;
    return (int)hipGetLastError();
}

} // extern "C"
