// uir_kernels.hip -- reference-compatible gfx950 kernels of libuir: the Joseph
// forward projector and backprojector, the elementwise kernels of
// ufo-ir-basic-ops.cl, finite differences, the smoothed-TV gradient and the
// reductions that ufo-ir ran as host loops.
//
// Built with -ffp-contract=off: every float operation rounds where the
// reference source rounds, so these kernels reproduce the reference's
// arithmetic operation for operation (sum order included).  The fast
// reduced-precision batched kernels live in uir_fast.hip.
#include <hip/hip_runtime.h>

#include <cstring>

#include "uir_device.h"

namespace {

constexpr int kBlockSamples = 64; // projector-parallel-joseph.cl:4

// --------------------------------------------------------------------------
// Joseph forward projection, one thread per ray (detector s, angle).
// --------------------------------------------------------------------------
struct FpRefArgs {
    const float *vol;
    const float *r_sino;
    float *w_sino;
    const float *ray_w;
    const float *sin_lut;
    const float *cos_lut;
    int n, n_dets, ang0, n_ang, dir;
    float axis, corr;
};

// Per-ray set-up of projector-parallel-joseph.cl:51-75 / :112-136: returns
// false for an inactive detector; else the start coordinate along the ray's
// interpolation axis (a double expression rounded to float, :73 / :134), the
// per-line step and the family (0 = FP_hor, 1 = FP_vert).
__device__ __forceinline__ bool fp_ray_setup(int s, int ang, int N, int D, float axis, const float *sin_lut,
                                             const float *cos_lut, int dir_in, float &pos, float &step, int &dir)
{
    const float required_width = axis * 2;
    const float diff = (float)D - required_width;
    if ((diff < 0 && (float)s < fabsf(diff)) || (diff > 0 && (float)s > required_width))
        return false;
    const float rot = diff / 2.0f;
    const float origin_shift = (float)N / 2.0f;
    const float st = sin_lut[ang], ct = cos_lut[ang];
    dir = dir_in >= 0 ? dir_in : (fabsf(st) <= fabsf(ct));
    const double t = ((0.5 + (double)rot) + (double)s) - (double)(0.5f * (float)D);
    if (dir == 0) {
        const float det_step = -1.0f / st;
        step = ct / st;
        pos = (float)(t * (double)det_step + (double)((-origin_shift) * step) + (double)(0.5f * (float)N));
    } else {
        const float det_step = 1.0f / ct;
        step = st / ct;
        pos = (float)(t * (double)det_step + (double)((-origin_shift) * step) + (double)(0.5f * (float)N));
    }
    return true;
}

// One block sum of :82-90: samples i in [b0, bend) of the ray, pos advanced
// by step per sample (the incremental float stepping of the reference).
// FP_hor walks column i linear in y; FP_vert walks row i linear in x.  Samples
// past the image edge read the zero border and do not change the sum.
// Samples are taken in batches of kFpBatch: the batch's coordinates follow
// the same sequential float additions, all its taps are loaded (clamped,
// always-valid addresses, zero where the reference reads the border), and
// only then are the samples added to the block sum in order -- the
// arithmetic and its order are the reference's, the loads overlap.
constexpr int kFpBatch = 8;

__device__ __forceinline__ float fp_block_sum(const float *img, int N, int dir, int b0, int bend, float &pos, float step)
{
    float inner = 0.0f;
    for (int i0 = b0; i0 < bend; i0 += kFpBatch) {
        float w[kFpBatch], va[kFpBatch], vb[kFpBatch];
        int c0[kFpBatch];
#pragma unroll
        for (int k = 0; k < kFpBatch; ++k) {
            const int i = min(i0 + k, bend - 1); // lanes past bend load a valid address, unused
            const float v = pos - 0.5f;
            const float fv = floorf(v);
            c0[k] = (int)fv;
            w[k] = v - fv;
            const int ca = min(max(c0[k], 0), N - 1), cb = min(max(c0[k] + 1, 0), N - 1);
            // FP_hor walks column i linear in y; FP_vert walks row i linear in x
            const size_t oa = dir == 0 ? (size_t)ca * N + i : (size_t)i * N + ca;
            const size_t ob = dir == 0 ? (size_t)cb * N + i : (size_t)i * N + cb;
            va[k] = img[oa];
            vb[k] = img[ob];
            if (i0 + k < bend)
                pos += step;
        }
        // An empty statement that takes the loaded registers: every load of the
        // batch is issued before the first one is waited for (otherwise the
        // compiler moves a load whose value is only used under the bounds test
        // into a branch and waits on it there).
#pragma unroll
        for (int k = 0; k < kFpBatch; ++k)
            asm volatile("" : "+v"(va[k]), "+v"(vb[k]));
#pragma unroll
        for (int k = 0; k < kFpBatch; ++k) {
            if (i0 + k < bend) {
                const float t0 = (c0[k] >= 0 && c0[k] < N) ? va[k] : 0.0f;
                const float t1 = (c0[k] + 1 >= 0 && c0[k] + 1 < N) ? vb[k] : 0.0f;
                const float omw = 1.0f - w[k];
                inner += omw * t0 + w[k] * t1;
            }
        }
    }
    return inner;
}

// Inactive detector (:63-66): the reference kernel returns and the (aliased)
// sinogram keeps r; the SIRT / SART form then multiplies by the ray weight.
__device__ __forceinline__ void fp_inactive(const FpRefArgs &p, size_t o)
{
    if (p.ray_w)
        p.w_sino[o] = p.r_sino[o] * p.ray_w[o];
    else if (p.w_sino != p.r_sino)
        p.w_sino[o] = p.r_sino[o];
}

__device__ __forceinline__ void fp_epilogue(const FpRefArgs &p, size_t o, float acc)
{
    float w = p.r_sino[o] + acc * p.corr; // :92-94
    if (p.ray_w)
        w = w * p.ray_w[o];
    p.w_sino[o] = w;
}

__global__ __launch_bounds__(256) void k_fp_ref(FpRefArgs p)
{
    const int s = blockIdx.x * 64 + threadIdx.x;
    const int k = blockIdx.y * 4 + threadIdx.y;
    if (s >= p.n_dets || k >= p.n_ang)
        return;
    const int ang = p.ang0 + k;
    const int N = p.n;
    const size_t o = (size_t)ang * p.n_dets + s;
    float pos, step;
    int dir;
    if (!fp_ray_setup(s, ang, N, p.n_dets, p.axis, p.sin_lut, p.cos_lut, p.dir, pos, step, dir)) {
        fp_inactive(p, o);
        return;
    }
    float acc = 0.0f;
    for (int b0 = 0; b0 < N; b0 += kBlockSamples)
        acc += fp_block_sum(p.vol, N, dir, b0, min(N, b0 + kBlockSamples), pos, step);
    fp_epilogue(p, o, acc);
}

// The same projection with each ray split over 16 threads: thread slot q
// takes blocks q, q + 16, ...; it reaches a block's start coordinate by
// replaying the reference's float additions (no memory traffic), so every
// sample sees the coordinate the sequential walk would see.  The block sums
// meet in LDS and are added in block order, as the reference adds them.
// Used where few rays are projected at once (one angle per launch in SART).
constexpr int kSplitRays = 16, kSplitSlots = 16, kMaxBlocks = 16384 / kBlockSamples;

__global__ __launch_bounds__(256) void k_fp_ref_split(FpRefArgs p)
{
    __shared__ float sums[kSplitRays][kMaxBlocks];
    const int r = threadIdx.x % kSplitRays, q = threadIdx.x / kSplitRays;
    const int s = blockIdx.x * kSplitRays + r;
    const int ang = p.ang0 + blockIdx.y;
    const int N = p.n;
    const int n_blocks = (N + kBlockSamples - 1) / kBlockSamples;
    float pos = 0.0f, step = 0.0f;
    int dir = 0;
    const bool live = s < p.n_dets &&
                      fp_ray_setup(s, ang, N, p.n_dets, p.axis, p.sin_lut, p.cos_lut, p.dir, pos, step, dir);
    if (live) {
        int at = 0; // blocks already stepped over (every block before the last is whole)
        for (int j = q; j < n_blocks; j += kSplitSlots) {
            for (; at < j; ++at) {
#pragma unroll
                for (int k = 0; k < kBlockSamples; ++k)
                    pos += step;
            }
            const int b0 = j * kBlockSamples;
            sums[r][j] = fp_block_sum(p.vol, N, dir, b0, min(N, b0 + kBlockSamples), pos, step);
            at = j + 1;
        }
    }
    __syncthreads();
    if (q != 0 || s >= p.n_dets)
        return;
    const size_t o = (size_t)ang * p.n_dets + s;
    if (!live) {
        fp_inactive(p, o);
        return;
    }
    float acc = 0.0f;
    for (int j = 0; j < n_blocks; ++j)
        acc += sums[r][j];
    fp_epilogue(p, o, acc);
}

// --------------------------------------------------------------------------
// The same projection with one wave per ray and the block starts from a table.
// --------------------------------------------------------------------------
// A ray's coordinate before sample i is the reference's sequential float walk
// (:82-90): i additions of the step to the start coordinate.  That walk does
// not depend on the image, so it is taken once per geometry: the table holds,
// for every ray, the coordinate at the start of each 64-sample block, reached
// by the same additions (k_fp_block_starts).  A ray's 64-sample blocks then
// spread over a wave, 16 samples per lane: a lane replays at most 48 additions
// from its block's start, loads all 32 taps of its samples at once, and the
// four lanes of a block add their samples into the block sum one after another
// in sample order (the value passes lane to lane); the block sums are added in
// block order.  Every sample sees the coordinate the sequential walk sees and
// every sum is taken in the reference's order: bit-identical to k_fp_ref.
constexpr int kSegSamples = 16; // samples per lane; 4 lanes per 64-sample block

__global__ __launch_bounds__(256) void k_fp_block_starts(FpRefArgs p, float *__restrict__ starts)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int ang = p.ang0 + blockIdx.y;
    if (s >= p.n_dets)
        return;
    const int N = p.n, n_blocks = (N + kBlockSamples - 1) / kBlockSamples;
    float pos = 0.0f, step = 0.0f;
    int dir = 0;
    float *out = starts + ((size_t)ang * p.n_dets + s) * n_blocks;
    if (!fp_ray_setup(s, ang, N, p.n_dets, p.axis, p.sin_lut, p.cos_lut, p.dir, pos, step, dir)) {
        for (int j = 0; j < n_blocks; ++j)
            out[j] = 0.0f;
        return;
    }
    for (int j = 0; j < n_blocks; ++j) {
        out[j] = pos;
#pragma unroll
        for (int k = 0; k < kBlockSamples; ++k)
            pos += step;
    }
}

__global__ __launch_bounds__(256) void k_fp_ref_seg(FpRefArgs p, const float *__restrict__ starts)
{
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6); // one ray per wave
    const int ang = p.ang0 + blockIdx.y;
    if (s >= p.n_dets)
        return; // uniform over the wave
    const int N = p.n, n_blocks = (N + kBlockSamples - 1) / kBlockSamples;
    float pos0 = 0.0f, step = 0.0f;
    int dir = 0;
    const size_t o = (size_t)ang * p.n_dets + s;
    if (!fp_ray_setup(s, ang, N, p.n_dets, p.axis, p.sin_lut, p.cos_lut, p.dir, pos0, step, dir)) {
        if (lane == 0)
            fp_inactive(p, o);
        return;
    }
    const float *st = starts + o * n_blocks;
    const int q = lane & 3; // this lane's quarter of its block
    // the epilogue's operands, loaded while the samples are taken
    const float r_in = lane == 0 ? p.r_sino[o] : 0.0f;
    const float rw_in = (lane == 0 && p.ray_w) ? p.ray_w[o] : 0.0f;
    float acc = 0.0f;
    for (int jb0 = 0; jb0 < n_blocks; jb0 += 16) {
        const int jb = jb0 + (lane >> 2);
        float pos = jb < n_blocks ? st[jb] : 0.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            if (r < q) {
#pragma unroll
                for (int k = 0; k < kSegSamples; ++k)
                    pos += step;
            }
        }
        const int i0 = jb * kBlockSamples + kSegSamples * q;
        float va[kSegSamples], vb[kSegSamples], w[kSegSamples];
        int c0[kSegSamples];
#pragma unroll
        for (int k = 0; k < kSegSamples; ++k) {
            const int i = min(i0 + k, N - 1); // lanes past the image load a valid address, unused
            const float v = pos - 0.5f;
            const float fv = floorf(v);
            c0[k] = (int)fv;
            w[k] = v - fv;
            const int ca = min(max(c0[k], 0), N - 1), cb = min(max(c0[k] + 1, 0), N - 1);
            const size_t oa = dir == 0 ? (size_t)ca * N + i : (size_t)i * N + ca;
            const size_t ob = dir == 0 ? (size_t)cb * N + i : (size_t)i * N + cb;
            va[k] = p.vol[oa];
            vb[k] = p.vol[ob];
            pos += step;
        }
#pragma unroll
        for (int k = 0; k < kSegSamples; ++k)
            asm volatile("" : "+v"(va[k]), "+v"(vb[k]));
        float val[kSegSamples];
#pragma unroll
        for (int k = 0; k < kSegSamples; ++k) {
            const float t0 = (c0[k] >= 0 && c0[k] < N) ? va[k] : 0.0f;
            const float t1 = (c0[k] + 1 >= 0 && c0[k] + 1 < N) ? vb[k] : 0.0f;
            const float omw = 1.0f - w[k];
            val[k] = omw * t0 + w[k] * t1;
        }
        // the block sum, lane q continuing from lane q - 1 (:84-89 order);
        // the hand-over is a DPP row shift by one lane (lanes of a block
        // share a row of 16)
        float run = 0.0f;
#pragma unroll
        for (int stage = 0; stage < 4; ++stage) {
            const float from = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(run), 0x111, 0xf, 0xf, false));
            if (q == stage) {
                float a = stage == 0 ? 0.0f : from;
#pragma unroll
                for (int k = 0; k < kSegSamples; ++k)
                    if (i0 + k < N)
                        a = a + val[k];
                run = a;
            }
        }
        // block sums in block order (:89)
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            const float bs = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(run), 4 * jj + 3));
            if (jb0 + jj < n_blocks)
                acc = acc + bs;
        }
    }
    if (lane == 0) { // fp_epilogue with the prefetched operands (:92-94)
        float w = r_in + acc * p.corr;
        if (p.ray_w)
            w = w * rw_in;
        p.w_sino[o] = w;
    }
}

// --------------------------------------------------------------------------
// Backprojection, one thread per pixel, per-run accumulation.
// --------------------------------------------------------------------------
struct BpRefArgs {
    const float *r_vol;
    float *w_vol;
    const float *sino;
    const float *pix_w;
    const float *sin_lut;
    const float *cos_lut;
    int n, n_dets, n_angles, posc;
    float axis, relax;
    uirk_runs runs;
};

__global__ __launch_bounds__(256) void k_bp_ref(BpRefArgs p)
{
    const int px = blockIdx.x * 64 + threadIdx.x;
    const int py = blockIdx.y * 4 + threadIdx.y;
    const int N = p.n, D = p.n_dets;
    if (px >= N || py >= N)
        return;
    const size_t o = (size_t)py * N + px;

    // projector-parallel-joseph.cl:174-190 (detector quantities use n_dets)
    const float required_width = p.axis * 2;
    const float diff = (float)D - required_width;
    const float half_active = diff < 0 ? (float)D - p.axis : p.axis;
    const float edge0 = p.axis - half_active + 0.5f;
    const float edge1 = p.axis + half_active - 0.5f;
    const float rot = diff / 2.0f;
    const float det_shift = (float)D / 2.0f;
    const float origin_shift = (float)N / 2.0f;
    const float shift = det_shift - rot;
    const float fX = (float)px + 0.5f - origin_shift;
    const float fY = (float)py + 0.5f - origin_shift;

    float out = p.pix_w ? 0.0f : p.r_vol[o];
    for (int r = 0; r < p.runs.count; ++r) {
        const int off = (int)p.runs.offset[r], cnt = (int)p.runs.n[r];
        float value = 0.0f;
        for (int k = 0; k < cnt; ++k) { // :196-214
            const int ang = off + k;
            const float st = p.sin_lut[ang], ct = p.cos_lut[ang];
            float u = fX * ct - fY * st + shift;
            if (u > edge1)
                u = edge1;
            else if (u < edge0)
                u = edge0;
            const float w = u - 0.5f;
            const float fw = floorf(w);
            const int i0 = (int)fw;
            const float a = w - fw;
            const float oma = 1.0f - a;
            const float *row = p.sino + (size_t)ang * D;
            const float t0 = (i0 >= 0 && i0 < D) ? row[i0] : 0.0f;
            const float t1 = (i0 + 1 >= 0 && i0 + 1 < D) ? row[i0 + 1] : 0.0f;
            value += oma * t0 + a * t1;
        }
        out = out + p.relax * value; // :216
    }
    if (p.pix_w) {
        out = out * p.pix_w[o];
        out = out + p.r_vol[o];
        if (p.posc)
            out = out > 0 ? out : 0.0f;
    }
    p.w_vol[o] = out;
}

// --------------------------------------------------------------------------
// Backprojection of one angle, four pixels per thread (SART's per-angle BP).
// --------------------------------------------------------------------------
// The arithmetic of k_bp_ref with one run of one angle (:196-216) in
// accumulate mode: w = r + relax * (0 + lerp).  Row-aligned buffers (N a
// multiple of 4, 16-byte aligned pointers) move four pixels per load and
// store.
template <bool VEC>
__global__ __launch_bounds__(256) void k_bp_ref_angle(BpRefArgs p, int ang)
{
    const int px0 = (blockIdx.x * 64 + threadIdx.x) * 4;
    const int py = blockIdx.y * 4 + threadIdx.y;
    const int N = p.n, D = p.n_dets;
    if (px0 >= N || py >= N)
        return;
    // projector-parallel-joseph.cl:174-190 (detector quantities use n_dets)
    const float required_width = p.axis * 2;
    const float diff = (float)D - required_width;
    const float half_active = diff < 0 ? (float)D - p.axis : p.axis;
    const float edge0 = p.axis - half_active + 0.5f;
    const float edge1 = p.axis + half_active - 0.5f;
    const float rot = diff / 2.0f;
    const float det_shift = (float)D / 2.0f;
    const float origin_shift = (float)N / 2.0f;
    const float shift = det_shift - rot;
    const float fY = (float)py + 0.5f - origin_shift;
    const float st = p.sin_lut[ang], ct = p.cos_lut[ang];
    const float *row = p.sino + (size_t)ang * D;
    const size_t o = (size_t)py * N + px0;
    float in[4];
    if (VEC) {
        const float4 v = *(const float4 *)(p.r_vol + o);
        in[0] = v.x, in[1] = v.y, in[2] = v.z, in[3] = v.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            in[k] = px0 + k < N ? p.r_vol[o + k] : 0.0f;
    }
    float out[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float fX = (float)(px0 + k) + 0.5f - origin_shift;
        float u = fX * ct - fY * st + shift;
        if (u > edge1)
            u = edge1;
        else if (u < edge0)
            u = edge0;
        const float w = u - 0.5f;
        const float fw = floorf(w);
        const int i0 = (int)fw;
        const float a = w - fw;
        const float oma = 1.0f - a;
        const float t0 = (i0 >= 0 && i0 < D) ? row[i0] : 0.0f;
        const float t1 = (i0 + 1 >= 0 && i0 + 1 < D) ? row[i0 + 1] : 0.0f;
        float value = 0.0f;
        value += oma * t0 + a * t1;
        out[k] = in[k] + p.relax * value; // :216
    }
    if (VEC) {
        *(float4 *)(p.w_vol + o) = make_float4(out[0], out[1], out[2], out[3]);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (px0 + k < N)
                p.w_vol[o + k] = out[k];
    }
}

// --------------------------------------------------------------------------
// Elementwise
// --------------------------------------------------------------------------
inline unsigned grid1d(size_t n, unsigned block = 256)
{
    size_t g = (n + block - 1) / block;
    if (g > 65536u * 16u)
        g = 65536u * 16u;
    return (unsigned)(g ? g : 1);
}

__global__ void k_set(float *out, size_t n, float v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = v;
}

__global__ void k_copy(float *dst, const float *src, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_inv(const float *in, float *out, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = in[i];
        out[i] = (v != 0) ? 1.0f / v : 0.0f; // ufo-ir-basic-ops.cl:34
    }
}

// Ray weights with a floor: rays whose FP(1) is below vmin are treated as
// missing the image (weight 0).
__global__ void k_inv_min(const float *in, float *out, size_t n, float vmin)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = in[i];
        out[i] = (v != 0 && v >= vmin) ? 1.0f / v : 0.0f;
    }
}

__global__ void k_binary(int op, const float *a, const float *b, float m, float *out, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float x = a[i], y = b[i];
        float r;
        switch (op) {
        case UIRK_MUL: r = x * y; break;
        case UIRK_ADD: r = x + y; break;
        case UIRK_SUB: r = x - y; break;
        case UIRK_SUB_SCALED: r = x - m * y; break;
        case UIRK_ADD_SCALED: r = x + m * y; break;
        case UIRK_MAX: r = fmaxf(x, y); break;
        default: r = x / y; break;
        }
        out[i] = r;
    }
}

__global__ void k_unary(int op, const float *in, float *out, size_t n, float m)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float x = in[i];
        float r;
        switch (op) {
        case UIRK_POSC: r = x > 0 ? x : 0.0f; break;
        case UIRK_SQRT: r = sqrtf(x); break;
        default: r = x * m; break;
        }
        out[i] = r;
    }
}

__global__ void k_mul_rows(const float *a, const float *b, float *out, uint32_t width, uint32_t offset, uint32_t rows)
{
    const size_t total = (size_t)width * rows;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t idx = (size_t)offset * width + i;
        out[idx] = a[idx] * b[idx];
    }
}

// --------------------------------------------------------------------------
// Reductions (deterministic: fixed grid, fixed tree order)
// --------------------------------------------------------------------------
constexpr int kRedBlock = 256;
constexpr unsigned kRedMaxBlocks = 1024;

inline unsigned red_blocks(size_t n)
{
    size_t g = (n + kRedBlock * 4 - 1) / (kRedBlock * 4);
    if (g > kRedMaxBlocks)
        g = kRedMaxBlocks;
    return (unsigned)(g ? g : 1);
}

template <typename T>
__device__ T block_sum(T v, T *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = kRedBlock / 2; w > 0; w >>= 1) {
        if (t < w)
            sh[t] = sh[t] + sh[t + w];
        __syncthreads();
    }
    return sh[0];
}

template <int KIND> // 0: sum |a|, 1: sum a*b
__global__ __launch_bounds__(kRedBlock) void k_red_f64(const float *a, const float *b, size_t n, double *ws)
{
    __shared__ double sh[kRedBlock];
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)kRedBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kRedBlock) {
        if (KIND == 0)
            acc += fabs((double)a[i]);
        else
            acc += (double)a[i] * (double)b[i];
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0)
        ws[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedBlock) void k_red_final_f64(const double *ws, unsigned nb, double *result)
{
    __shared__ double sh[kRedBlock];
    double acc = 0.0;
    for (unsigned i = threadIdx.x; i < nb; i += kRedBlock)
        acc += ws[i];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0)
        *result = s;
}

// ufo-ir-basic-ops-processor.c:181-193: part p sums a[p*i] * b[p*i] for
// i < part_len in float; one thread per part, in order.
__global__ void k_dot_compat_parts(const float *a, const float *b, uint32_t parts, uint32_t part_len, float *ws)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= parts)
        return;
    float part = 0.0f;
    for (uint32_t i = 0; i < part_len; ++i) {
        const size_t idx = (size_t)p * i;
        part += a[idx] * b[idx];
    }
    ws[p] = part;
}

__global__ void k_dot_compat_final(const float *ws, uint32_t parts, float *result)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float norm = 0.0f;
        for (uint32_t p = 0; p < parts; ++p)
            norm += ws[p];
        *result = norm;
    }
}

__global__ __launch_bounds__(kRedBlock) void k_minmax(const float *v, size_t n, float *ws)
{
    __shared__ float smx[kRedBlock], smn[kRedBlock];
    float mx = -3.40282347e+38f, mn = 3.40282347e+38f;
    for (size_t i = blockIdx.x * (size_t)kRedBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kRedBlock) {
        mx = fmaxf(v[i], mx);
        mn = fminf(v[i], mn);
    }
    const int t = threadIdx.x;
    smx[t] = mx;
    smn[t] = mn;
    __syncthreads();
    for (int w = kRedBlock / 2; w > 0; w >>= 1) {
        if (t < w) {
            smx[t] = fmaxf(smx[t], smx[t + w]);
            smn[t] = fminf(smn[t], smn[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        ws[2 * blockIdx.x] = smn[0];
        ws[2 * blockIdx.x + 1] = smx[0];
    }
}

__global__ void k_minmax_final(const float *ws, unsigned nb, float *minmax)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float mn = 3.40282347e+38f, mx = -3.40282347e+38f;
        for (unsigned i = 0; i < nb; ++i) {
            mn = fminf(ws[2 * i], mn);
            mx = fmaxf(ws[2 * i + 1], mx);
        }
        minmax[0] = mn;
        minmax[1] = mx;
    }
}

// ufo-ir-basic-ops-processor.c:362-366
__global__ void k_normalize_apply(float *v, size_t n, const float *minmax)
{
    const float mn = minmax[0], mx = minmax[1];
    const float delta = 1 / (mx - mn);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = delta * v[i] - mn * delta;
}

// --------------------------------------------------------------------------
// Finite differences and TV gradient
// --------------------------------------------------------------------------
__global__ void k_diff(int kind, const float *in, float *out, int w, int h)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= w || y >= h)
        return;
    const int idx = y * w + x;
    const int last = w * h;
    float r;
    switch (kind) {
    case UIRK_DX: r = in[idx] - in[idx - 1 + (x == 0 ? w : 0)]; break;            // Dx :1-11
    case UIRK_DXT: r = in[idx] - in[idx + 1 - (x == w - 1 ? w : 0)]; break;       // Dxt :13-24
    case UIRK_DY: r = in[idx] - in[idx - w + (y == 0 ? last : 0)]; break;         // Dy :26-37
    default: r = in[idx] - in[idx + w - (y == h - 1 ? last : 0)]; break;          // Dyt :39-51
    }
    out[idx] = r;
}

__device__ inline float nb_read(const float *img, int w, int h, int x, int y)
{
    return (x >= 0 && x < w && y >= 0 && y < h) ? img[(size_t)y * w + x] : 0.0f;
}

// ufo-math-tvstd-method.cl:5-57
__global__ void k_l1_grad(const float *in, float *out, int w, int h)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= w || y >= h)
        return;
    const float eps = 0.001f;
    const float f3 = nb_read(in, w, h, x, y);
    const float f1 = nb_read(in, w, h, x, y - 1);
    const float f0 = nb_read(in, w, h, x - 1, y - 1);
    const float f2 = nb_read(in, w, h, x - 1, y);
    const float f4 = nb_read(in, w, h, x + 1, y);
    const float f6 = nb_read(in, w, h, x + 1, y + 1);
    const float f5 = nb_read(in, w, h, x, y + 1);
    const float df = (2.0f * f3 - f2 - f5) / sqrtf(eps + (f3 - f2) * (f3 - f2) + (f3 - f5) * (f3 - f5))
                   - (f4 - f3) / sqrtf(eps + (f4 - f3) * (f4 - f3) + (f4 - f6) * (f4 - f6))
                   - (f1 - f3) / sqrtf(eps + (f1 - f3) * (f1 - f3) + (f1 - f0) * (f1 - f0));
    out[(size_t)y * w + x] = df;
}

} // namespace

// ==========================================================================
// Launchers (C ABI, see uir_device.h)
// ==========================================================================
extern "C" {

int uirk_fp_ref(const float *vol, const float *r_sino, float *w_sino, const float *ray_w,
                const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis,
                float corr, int ang0, int n_ang, int dir, void *stream)
{
    if (n_ang <= 0 || n_dets <= 0)
        return 0;
    FpRefArgs a{vol, r_sino, w_sino, ray_w, sin_lut, cos_lut, n, n_dets, ang0, n_ang, dir, axis, corr};
    dim3 block(64, 4), grid((n_dets + 63) / 64, (n_ang + 3) / 4);
    hipLaunchKernelGGL(k_fp_ref, grid, block, 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

int uirk_fp_ref_split(const float *vol, const float *r_sino, float *w_sino, const float *ray_w,
                      const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis,
                      float corr, int ang0, int n_ang, int dir, void *stream)
{
    if (n_ang <= 0 || n_dets <= 0)
        return 0;
    if (n > 16384)
        return (int)hipErrorInvalidValue;
    FpRefArgs a{vol, r_sino, w_sino, ray_w, sin_lut, cos_lut, n, n_dets, ang0, n_ang, dir, axis, corr};
    dim3 grid((n_dets + kSplitRays - 1) / kSplitRays, n_ang);
    hipLaunchKernelGGL(k_fp_ref_split, grid, dim3(256), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

int uirk_fp_block_starts(const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis, int n_angles,
                         int dir, float *starts, void *stream)
{
    if (n_angles <= 0 || n_dets <= 0 || n <= 0)
        return 0;
    FpRefArgs a{nullptr, nullptr, nullptr, nullptr, sin_lut, cos_lut, n, n_dets, 0, n_angles, dir, axis, 0.0f};
    dim3 grid((n_dets + 255) / 256, n_angles);
    hipLaunchKernelGGL(k_fp_block_starts, grid, dim3(256), 0, (hipStream_t)stream, a, starts);
    return (int)hipGetLastError();
}

int uirk_fp_ref_seg(const float *vol, const float *r_sino, float *w_sino, const float *ray_w,
                    const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis,
                    float corr, int ang0, int n_ang, int dir, const float *starts, void *stream)
{
    if (n_ang <= 0 || n_dets <= 0)
        return 0;
    FpRefArgs a{vol, r_sino, w_sino, ray_w, sin_lut, cos_lut, n, n_dets, ang0, n_ang, dir, axis, corr};
    dim3 grid((n_dets + 3) / 4, n_ang);
    hipLaunchKernelGGL(k_fp_ref_seg, grid, dim3(256), 0, (hipStream_t)stream, a, starts);
    return (int)hipGetLastError();
}

int uirk_bp_ref_angle(const float *r_vol, float *w_vol, const float *sino, const float *sin_lut,
                      const float *cos_lut, int n, int n_dets, float axis, float relax, int ang, void *stream)
{
    if (n <= 0)
        return 0;
    BpRefArgs a;
    memset(&a, 0, sizeof a);
    a.r_vol = r_vol;
    a.w_vol = w_vol;
    a.sino = sino;
    a.sin_lut = sin_lut;
    a.cos_lut = cos_lut;
    a.n = n;
    a.n_dets = n_dets;
    a.axis = axis;
    a.relax = relax;
    dim3 block(64, 4), grid((n + 255) / 256, (n + 3) / 4);
    const bool vec = (n & 3) == 0 && (((uintptr_t)r_vol | (uintptr_t)w_vol) & 15) == 0;
    if (vec)
        hipLaunchKernelGGL(k_bp_ref_angle<true>, grid, block, 0, (hipStream_t)stream, a, ang);
    else
        hipLaunchKernelGGL(k_bp_ref_angle<false>, grid, block, 0, (hipStream_t)stream, a, ang);
    return (int)hipGetLastError();
}

int uirk_bp_ref(const float *r_vol, float *w_vol, const float *sino, int n_angles,
                const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis,
                float relax, const uirk_runs *runs, const float *pix_w, int posc, void *stream)
{
    if (n <= 0)
        return 0;
    BpRefArgs a;
    a.r_vol = r_vol;
    a.w_vol = w_vol;
    a.sino = sino;
    a.pix_w = pix_w;
    a.sin_lut = sin_lut;
    a.cos_lut = cos_lut;
    a.n = n;
    a.n_dets = n_dets;
    a.n_angles = n_angles;
    a.posc = posc;
    a.axis = axis;
    a.relax = relax;
    a.runs = *runs;
    dim3 block(64, 4), grid((n + 63) / 64, (n + 3) / 4);
    hipLaunchKernelGGL(k_bp_ref, grid, block, 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

int uirk_set(float *out, size_t n, float v, void *stream)
{
    if (!n)
        return 0;
    hipLaunchKernelGGL(k_set, dim3(grid1d(n)), dim3(256), 0, (hipStream_t)stream, out, n, v);
    return (int)hipGetLastError();
}

int uirk_copy(float *dst, const float *src, size_t n, void *stream)
{
    if (!n)
        return 0;
    hipLaunchKernelGGL(k_copy, dim3(grid1d(n)), dim3(256), 0, (hipStream_t)stream, dst, src, n);
    return (int)hipGetLastError();
}

int uirk_inv_min(const float *in, float *out, size_t n, float vmin, void *stream)
{
    if (!n)
        return 0;
    hipLaunchKernelGGL(k_inv_min, dim3(grid1d(n)), dim3(256), 0, (hipStream_t)stream, in, out, n, vmin);
    return (int)hipGetLastError();
}

int uirk_inv(const float *in, float *out, size_t n, void *stream)
{
    if (!n)
        return 0;
    hipLaunchKernelGGL(k_inv, dim3(grid1d(n)), dim3(256), 0, (hipStream_t)stream, in, out, n);
    return (int)hipGetLastError();
}

int uirk_binary(int op, const float *a, const float *b, float m, float *out, size_t n, void *stream)
{
    if (!n)
        return 0;
    hipLaunchKernelGGL(k_binary, dim3(grid1d(n)), dim3(256), 0, (hipStream_t)stream, op, a, b, m, out, n);
    return (int)hipGetLastError();
}

int uirk_unary(int op, const float *in, float *out, size_t n, float m, void *stream)
{
    if (!n)
        return 0;
    hipLaunchKernelGGL(k_unary, dim3(grid1d(n)), dim3(256), 0, (hipStream_t)stream, op, in, out, n, m);
    return (int)hipGetLastError();
}

int uirk_mul_rows(const float *a, const float *b, float *out, uint32_t width, uint32_t offset,
                  uint32_t rows, void *stream)
{
    const size_t n = (size_t)width * rows;
    if (!n)
        return 0;
    hipLaunchKernelGGL(k_mul_rows, dim3(grid1d(n)), dim3(256), 0, (hipStream_t)stream, a, b, out, width, offset, rows);
    return (int)hipGetLastError();
}

size_t uirk_reduce_ws_bytes(size_t n)
{
    // f64 partials (kRedMaxBlocks), minmax pairs, compat parts (<= n)
    size_t a = kRedMaxBlocks * sizeof(double) * 2;
    size_t b = n * sizeof(float) + 64;
    return a > b ? a : b;
}

int uirk_l1_f64(const float *v, size_t n, double *ws, double *result_dev, void *stream)
{
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL(k_red_f64<0>, dim3(nb), dim3(kRedBlock), 0, (hipStream_t)stream, v, v, n, ws);
    hipLaunchKernelGGL(k_red_final_f64, dim3(1), dim3(kRedBlock), 0, (hipStream_t)stream, ws, nb, result_dev);
    return (int)hipGetLastError();
}

int uirk_dot_f64(const float *a, const float *b, size_t n, double *ws, double *result_dev, void *stream)
{
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL(k_red_f64<1>, dim3(nb), dim3(kRedBlock), 0, (hipStream_t)stream, a, b, n, ws);
    hipLaunchKernelGGL(k_red_final_f64, dim3(1), dim3(kRedBlock), 0, (hipStream_t)stream, ws, nb, result_dev);
    return (int)hipGetLastError();
}

int uirk_dot_compat(const float *a, const float *b, uint32_t w, uint32_t h, float *ws,
                    float *result_dev, void *stream)
{
    if (h == 0)
        return (int)hipErrorInvalidValue;
    const uint32_t part_len = (uint32_t)(((size_t)w * h) / h);
    hipLaunchKernelGGL(k_dot_compat_parts, dim3((h + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       a, b, h, part_len, ws);
    hipLaunchKernelGGL(k_dot_compat_final, dim3(1), dim3(64), 0, (hipStream_t)stream, ws, h, result_dev);
    return (int)hipGetLastError();
}

int uirk_minmax(const float *v, size_t n, float *ws, float *minmax_dev, void *stream)
{
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL(k_minmax, dim3(nb), dim3(kRedBlock), 0, (hipStream_t)stream, v, n, ws);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(64), 0, (hipStream_t)stream, ws, nb, minmax_dev);
    return (int)hipGetLastError();
}

int uirk_normalize_apply(float *v, size_t n, const float *minmax_dev, void *stream)
{
    if (!n)
        return 0;
    hipLaunchKernelGGL(k_normalize_apply, dim3(grid1d(n)), dim3(256), 0, (hipStream_t)stream, v, n, minmax_dev);
    return (int)hipGetLastError();
}

int uirk_diff(int kind, const float *in, float *out, uint32_t w, uint32_t h, void *stream)
{
    if (!w || !h)
        return 0;
    dim3 block(64, 4), grid((w + 63) / 64, (h + 3) / 4);
    hipLaunchKernelGGL(k_diff, grid, block, 0, (hipStream_t)stream, kind, in, out, (int)w, (int)h);
    return (int)hipGetLastError();
}

int uirk_l1_grad(const float *in, float *out, uint32_t w, uint32_t h, void *stream)
{
    if (!w || !h)
        return 0;
    dim3 block(64, 4), grid((w + 63) / 64, (h + 3) / 4);
    hipLaunchKernelGGL(k_l1_grad, grid, block, 0, (hipStream_t)stream, in, out, (int)w, (int)h);
    return (int)hipGetLastError();
}

} // extern "C"

// ==========================================================================
// Allocation-free strided "dot" (uir_op_dot_compat_noalloc)
// ==========================================================================
// The same two kernels as uir_op_dot_compat (one float partial per row, then
// the partials added in row order), with the row partials in this file-scope
// device array instead of a workspace the library allocates, and the result
// written to the caller's device memory without a host synchronisation.  Each
// device holds its own copy of the array for the life of the process.  Every
// partial the final kernel reads is written by the same call's first kernel,
// so nothing needs clearing between calls; calls that overlap on one device
// race on the array.
constexpr uint32_t kNoAllocMaxRows = 65536; // 256 KiB of FP32 partials
static __device__ __align__(256) float scratch[kNoAllocMaxRows];

#include "uir_internal.h"

namespace {
constexpr int kMaxDevices = 64;
float *g_scratch_dev[kMaxDevices]; // the array's device address, per device
pthread_mutex_t g_scratch_lock = PTHREAD_MUTEX_INITIALIZER;

// The device address of `scratch` on the current device, looked up once per
// device (the host-side name is not a device address).
uir_status scratch_address(float **out)
{
    int dev = -1;
    uir_status s = uir_check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (s != UIR_OK)
        return s;
    if (dev < 0 || dev >= kMaxDevices)
        return uir_set_error(UIR_ERR_UNSUPPORTED, "device %d: at most %d devices", dev, kMaxDevices);
    pthread_mutex_lock(&g_scratch_lock);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        s = uir_check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (s == UIR_OK && !p)
            s = uir_set_error(UIR_ERR_DEVICE, "hipGetSymbolAddress(scratch) gave NULL");
        if (s == UIR_OK)
            g_scratch_dev[dev] = (float *)p;
    }
    *out = g_scratch_dev[dev];
    pthread_mutex_unlock(&g_scratch_lock);
    return s;
}
} // namespace

extern "C" uir_status uir_op_dot_compat_noalloc(const float *a, const float *b, uint32_t width, uint32_t height,
                                                float *d_result, void *stream)
{
    if (!a || !b || !d_result)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (height == 0)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "height is 0");
    if (height > kNoAllocMaxRows)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT,
                             "height %u exceeds the %u rows the allocation-free dot holds (uir_op_dot_compat "
                             "takes any height)", height, kNoAllocMaxRows);
    float *ws = nullptr;
    uir_status s = scratch_address(&ws);
    if (s != UIR_OK)
        return s;
    const uint32_t part_len = (uint32_t)(((size_t)width * height) / height); // ufo-ir-basic-ops-processor.c:181-183
    hipLaunchKernelGGL(k_dot_compat_parts, dim3((height + 255) / 256), dim3(256), 0, (hipStream_t)stream, a, b,
                       height, part_len, ws);
    hipLaunchKernelGGL(k_dot_compat_final, dim3(1), dim3(64), 0, (hipStream_t)stream, (const float *)ws, height,
                       d_result);
    return uir_check_hip(hipGetLastError(), "dot_compat launch");
}
