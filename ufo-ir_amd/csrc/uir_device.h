/*
 * uir_device.h -- internal C interface between the C host layer and the HIP
 * kernels of libuir (not installed; the public boundary is include/uir.h).
 *
 * Every launcher enqueues on the given hipStream_t (passed as void *) and
 * returns a hipError_t value (0 = success).
 */
#ifndef UIR_DEVICE_H
#define UIR_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#include "uir.h"

#ifdef __cplusplus
extern "C" {
#endif

#define UIRK_MAX_RUNS 16

typedef struct {
    uint32_t offset[UIRK_MAX_RUNS];
    uint32_t n[UIRK_MAX_RUNS];
    int32_t count;
} uirk_runs;

/* Measured ceilings of the projectors' binding resources on the current
 * device (uir_probe.hip): LDS read bytes, wave64 FP32 FMAs and wave64
 * float-to-int conversions per clock per CU, at the device's reported clock.  Returns 0, or 2 (no memory) / 3 (HIP
 * error). */
typedef struct {
    int n_cu;
    double clock_ghz;
    double lds_bytes_per_clock_cu;
    double fma_wave64_per_clock_cu;
    double cvt_wave64_per_clock_cu; /* v_cvt_flr_i32_f32 */
    double lds_ms, fma_ms; /* one launch of each loop */
} uirk_ceilings;
int uirk_probe_ceilings(void *stream, uirk_ceilings *c);

/* Reference-compatible Joseph FP (projector-parallel-joseph.cl:36-156) over
 * angles [ang0, ang0 + n_ang).  dir < 0 takes each angle's direction from the
 * LUT (|sin| <= |cos| -> vertical), else forces FP_vert (1) / FP_hor (0).
 * ray_w == NULL: w = r + corr * sum.  ray_w != NULL: w = (r + corr * sum) * ray_w
 * (the SIRT residual followed by operation_mul, ufo-ir-sirt-task.c:205-207). */
int uirk_fp_ref(const float *vol, const float *r_sino, float *w_sino, const float *ray_w,
                const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis,
                float corr, int ang0, int n_ang, int dir, void *stream);

/* The same FP, bit-identical, with each ray's 64-sample blocks spread over 16
 * threads (for launches over few rays, e.g. one angle); n <= 16384. */
int uirk_fp_ref_split(const float *vol, const float *r_sino, float *w_sino, const float *ray_w,
                      const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis,
                      float corr, int ang0, int n_ang, int dir, void *stream);

/* Reference-compatible BP (projector-parallel-joseph.cl:158-218) over a list of
 * subsets (runs), one per-run accumulation each: w = r + relax * sum_run.
 * pix_w == NULL: accumulate mode (r_vol may equal w_vol).
 * pix_w != NULL: SIRT mode, tmp = sum_runs(relax * S_run) (from 0), then
 * w = tmp * pix_w + r (ufo-ir-sirt-task.c:208-212); posc clamps the result. */
/* The coordinate at the start of every 64-sample block of every ray of
 * angles [0, n_angles) ([angle][detector][block], ceil(n / 64) blocks), taken
 * by the reference's sequential float walk; inactive rays get 0. */
int uirk_fp_block_starts(const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis, int n_angles,
                         int dir, float *starts, void *stream);
/* The same FP as uirk_fp_ref, bit-identical, one wave per ray, from the
 * block-start table of uirk_fp_block_starts (built for the same geometry,
 * direction rule and LUTs). */
int uirk_fp_ref_seg(const float *vol, const float *r_sino, float *w_sino, const float *ray_w,
                    const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis,
                    float corr, int ang0, int n_ang, int dir, const float *starts, void *stream);
/* BP of one angle in accumulate mode (w = r + relax * BP_ang), bit-identical
 * to uirk_bp_ref with a one-angle run. */
int uirk_bp_ref_angle(const float *r_vol, float *w_vol, const float *sino, const float *sin_lut,
                      const float *cos_lut, int n, int n_dets, float axis, float relax, int ang, void *stream);

int uirk_bp_ref(const float *r_vol, float *w_vol, const float *sino, int n_angles,
                const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis,
                float relax, const uirk_runs *runs, const float *pix_w, int posc, void *stream);

/* ---- The line projection model (uir_line.hip) ----
 * Ray s of angle a is the line X c - Y s_ = t, t = s + 1/2 - axis, with (c, s_)
 * the angle's LUT pair scaled to unit length and a component below 2^-20
 * taken as 0; its weight on a pixel is the length of that line inside the
 * pixel's unit square (centre (x + 1/2 - n/2, y + 1/2 - n/2)).  With d the
 * ray's offset from the pixel centre, M = max(|c|, |s_|) and m = min, the
 * length is (1/M) clamp(((M + m)/2 - |d|) / m, 0, 1); for m = 0 the pixel
 * takes d in [-1/2, 1/2).  Per sample of a ray (one line of pixels along
 * its major axis) and per angle of a pixel the model has two taps at
 * fractional position f in [0, 1) between them, with weights linear in f.
 * Constants per angle and geometry, computed on the host in double. */
typedef struct {
    double fp_base, fp_tk; /* FP: position of detector s at sample i: v = fp_base + s fp_tk + i fp_step */
    float fp_step;
    float fp_k, fp_b0, fp_b1; /* FP taps: w0 = clamp(fp_b0 - fp_k f), w1 = clamp(fp_b1 + fp_k f), x 1/M */
    float hm;                 /* 1/M */
    float c, s;               /* unit direction */
    float bp_off;             /* BP: u = x c - Y s_ + bp_off, bins floor(u), floor(u) + 1 */
    float bp_k, bp_b0, bp_b1; /* BP taps: w0 = clamp(bp_b0 - bp_k f, 0, 1/M), w1 = clamp(bp_b1 + bp_k f, 0, 1/M) */
    int32_t family;           /* 0: |c| >= |s_|, a sample per image row (taps x, x + 1); 1: per column (y, y + 1) */
    int32_t aligned;          /* m == 0: one tap, by the half-open rule */
    int32_t tie0;             /* aligned FP: f == 1/2 belongs to tap 0 */
} uirk_line_angle;

/* SIRT pixel weights of the line model: C = 1 / BP(1) where BP(1) >= this,
 * else 0 (a thousandth of a pixel's chord).  A pixel that rays only touch at
 * a corner has BP(1) = 0 exactly and about 1e-7 in FP32; without the floor it
 * would get C ~ 1e7 and the full residual of those rays. */
#define UIRK_LINE_PIX_MIN (1.0f / 1024.0f)

/* FP of angles [a0, a0 + na), all of one family: w = (r + corr FP) (* ray_w
 * when ray_w != NULL), r_sino may equal w_sino.  few_rays selects the layout
 * for launches of few rays (one angle): each ray spread over 16 lanes. */
int uirk_fp_line(const float *vol, const float *r_sino, float *w_sino, const float *ray_w,
                 const uirk_line_angle *d_tab, int family, int n, int n_dets, int a0, int na, float corr,
                 int few_rays, void *stream);
/* BP of angles [a0, a0 + na): pix_w == NULL: w_vol = r_vol + relax BP;
 * else w_vol = (relax BP) pix_w + r_vol, clamped at 0 when posc. */
int uirk_bp_line(const float *r_vol, float *w_vol, const float *sino, const float *pix_w, int posc,
                 const uirk_line_angle *d_tab, int n, int n_dets, int a0, int na, float relax, void *stream);

/* Elementwise (ufo-ir-basic-ops.cl). */
int uirk_set(float *out, size_t n, float v, void *stream);
int uirk_inv(const float *in, float *out, size_t n, void *stream);
/* out = (in != 0 && in >= vmin) ? 1 / in : 0 */
int uirk_inv_min(const float *in, float *out, size_t n, float vmin, void *stream);
int uirk_binary(int op, const float *a, const float *b, float m, float *out, size_t n, void *stream);
int uirk_mul_rows(const float *a, const float *b, float *out, uint32_t width, uint32_t offset,
                  uint32_t rows, void *stream);
int uirk_unary(int op, const float *in, float *out, size_t n, float m, void *stream);
int uirk_copy(float *dst, const float *src, size_t n, void *stream);

enum { UIRK_MUL = 0, UIRK_ADD = 1, UIRK_SUB = 2, UIRK_SUB_SCALED = 3, UIRK_ADD_SCALED = 4,
       UIRK_MAX = 5, UIRK_DIV = 6 };
enum { UIRK_POSC = 0, UIRK_SQRT = 1, UIRK_SCALE = 2 };

/* Reductions.  ws: device workspace of at least uirk_reduce_ws_bytes(n)
 * bytes; result_dev: one double (l1, dot) or float (compat) in device memory. */
size_t uirk_reduce_ws_bytes(size_t n);
int uirk_l1_f64(const float *v, size_t n, double *ws, double *result_dev, void *stream);
int uirk_dot_f64(const float *a, const float *b, size_t n, double *ws, double *result_dev, void *stream);
int uirk_dot_compat(const float *a, const float *b, uint32_t w, uint32_t h, float *ws,
                    float *result_dev, void *stream);
int uirk_minmax(const float *v, size_t n, float *ws, float *minmax_dev, void *stream);
int uirk_normalize_apply(float *v, size_t n, const float *minmax_dev, void *stream);

/* ---- Table-driven LDS-staged projectors (uir_fast.hip) ---- */

/* Per-slice descriptor.  Images are n x n, sinograms n_angles x n_dets, all in
 * the storage dtype of the launch except ray_w / pix_w (FP32). */
typedef struct {
    const void *x;       /* FP input image; BP reads the previous image here */
    void *x_out;         /* BP output image (may equal x) */
    const void *b;       /* FP input sinogram (r_sinogram / measured data) */
    void *r;             /* FP output sinogram (may equal b) */
    const float *ray_w;  /* FP: multiply the result by ray_w (NULL: none) */
    const float *pix_w;  /* ref BP: SIRT pixel weights (NULL: accumulate form) */
    const float *trig;   /* per angle (sin, cos) interleaved, FP32 */
    int32_t n, n_dets, n_angles, posc;
    float axis, corr, relax, pix_scale;
    int32_t n_runs;
    uint32_t run_off[UIRK_MAX_RUNS], run_n[UIRK_MAX_RUNS];
    /* fast BP only: x_out = x + relax * pix_scale * sum + tv_scale * tv
     * (tv: FP32 N x N, NULL = no term) */
    const float *tv;
    float tv_scale;
    int32_t pad_;
    /* line model (mode 2) only: 8 floats per angle, (fp_k, fp_b, 1/M, 0,
     * bp_k, bp_b0, bp_b1, 1/M) (uir_line_fast_consts); trig then holds the
     * model's unit directions.  With tv set, the BP is a split-Bregman sweep
     * x_out = x + relax P (mu sum + lambda tv), P = pix_w / (mu + 8 lambda
     * pix_w) per pixel (1 / (8 lambda) where pix_w = 0), with mu in pix_scale
     * and lambda in tv_scale. */
    const float *line;
} uirk_slice;

/* FP work item: angles [a0, a0 + na) of one direction family (0 = horizontal,
 * 1 = vertical) and detectors [d0, d0 + 64) of one slice; lines [l_lo, l_hi)
 * hold every sample of these rays that can touch the image (the other
 * samples read the zero border). */
typedef struct {
    int32_t slice, dir, a0, na, d0, l_lo, l_hi, pad;
} uirk_fp_item;

/* BP work item: the 64 x 64 pixel tile at (x0, y0) of one slice.  pair (mode
 * 3 only): the tile of slice + 1 too, which has the same geometry (the batch
 * sets it, not the planner). */
typedef struct {
    int32_t slice, x0, y0, pair;
} uirk_bp_item;

/* XCD-aware partition of a work table: workgroup g runs on XCD g % n_xcd
 * (round-robin dispatch) and processes only items [start[x], start[x + 1])
 * of its XCD x, so each XCD's L2 holds the data of a contiguous,
 * cost-balanced share of the table (a few slices, not all of them).  n_xcd <= 1
 * or a grid that is not a multiple of n_xcd falls back to plain striding. */
#define UIRK_MAX_XCD 16
typedef struct {
    int32_t n_xcd;
    int32_t start[UIRK_MAX_XCD + 1];
} uirk_xcd_part;

#define UIRK_FP_DETS 64     /* detectors per FP work item */
#define UIRK_FP_ANGLES 16   /* angles per FP work item (default shape) */
#define UIRK_FP_ANGLES_MAX 32 /* widest FP shape */
#define UIRK_FP_LINES 32    /* lines per staged chunk (divides the 64-sample block) */
#define UIRK_FP_WMAX 192    /* max footprint width of the pair-staged FP (a multiple of 64) */
#define UIRK_FP_WMAX_SINGLE 256 /* max footprint width of the single-value FP (a multiple of 64) */
#define UIRK_BP_TILE 64     /* BP tile edge */

/* Tile shapes of the fast path, chosen per geometry by the autotuner
 * (host/uir_tune.c): FP angles per work item (16, 8 or 32), BP angles per staged
 * batch (32 or 16) and BP tile rows (64 or 32; tiles are 64 pixels wide).
 * The reference-compatible mode takes the default {16, 32, 64} only. */
typedef struct {
    int32_t fp_angles, bp_angles, bp_rows;
    /* FP staging: 0 = (e, d) pairs (k_fp_pairs, windows up to UIRK_FP_WMAX),
     * 1 = single values (k_fp_tab, windows up to UIRK_FP_WMAX_SINGLE), for
     * geometries whose angle blocks the narrower window would cut short */
    int32_t fp_single;
} uirk_tile;
#define UIRK_TILE_DEFAULT {UIRK_FP_ANGLES, 32, UIRK_BP_TILE, 0}
/* Footprint width limit of a tile's FP. */
#define UIRK_TILE_FP_WMAX(t) ((t).fp_single ? UIRK_FP_WMAX_SINGLE : UIRK_FP_WMAX)

/* mode: 0 = reference-compatible FP32 arithmetic, 1 = fast path (FP32 math on
 * FP32 / FP16 / BF16 storage), 3 = (BP only) the fast path over slice pairs
 * (items with pair set cover slices slice and slice + 1 of one geometry, with
 * the same results as mode 1; d_slices required), 4 = (BP only) mode 2 over
 * slice pairs, 2 = the fast path with the line model (weights
 * from each slice's `line` table; the tile's bp_rows (64 or 32) only, the rest
 * default; BP: pix_w, when given, multiplies the update).  dtype: uir_dtype (FP32 / FP16 / BF16) of the
 * sinogram buffers (b, r); img_dtype: of the image buffers (x, x_out), either
 * dtype or FP32 (the reference-compatible mode takes FP32 only).  d_slices: device
 * descriptor array indexed by item.slice, or NULL to use *one (host memory,
 * passed by value as a kernel argument; item.slice must then be 0).
 * err: device int, set non-zero if a footprint exceeded the staging window. */
int uirk_fp_table(const uirk_slice *d_slices, const uirk_slice *one, const uirk_fp_item *d_items, int n_items,
                  const uirk_xcd_part *part, int mode, int dtype, int img_dtype, const uirk_tile *tile, int grid,
                  int *d_err, void *stream);
int uirk_bp_table(const uirk_slice *d_slices, const uirk_slice *one, const uirk_bp_item *d_items, int n_items,
                  const uirk_xcd_part *part, int mode, int dtype, int img_dtype, const uirk_tile *tile, int grid,
                  int *d_err, void *stream);
/* Ray weights of the fast path: ray_w = 1 / FP(1) with FP(1) evaluated in FP64
 * (exact coordinates, (sin, cos) from trig); 0 where FP(1) = 0 or < r_min. */
int uirk_ray_weights64(const float *d_trig, int n, int n_dets, int n_angles, float axis, float r_min, float *d_ray_w,
                       void *stream);
/* Work-table construction (host/uir_plan.c). */
typedef struct {
    uirk_fp_item *v;
    size_t n, cap;
} uirk_fp_vec;
typedef struct {
    uirk_bp_item *v;
    size_t n, cap;
} uirk_bp_vec;
/* Appends the FP items of one slice for the given subsets (runs or single
 * angles); angle blocks are sized so that each footprint fits the staging
 * window of wmax positions (UIRK_TILE_FP_WMAX of the tile).  Returns 0 on
 * success, -1 on allocation failure. */
int uirk_plan_fp(uirk_fp_vec *out, int slice, int n, int n_dets, float axis, const float *sin_lut,
                 const float *cos_lut, const uir_subset *subs, uint32_t n_subs, int max_angles, int wmax);
/* The same for the line model: trig pairs are its unit directions and ray
 * positions lie half a line step further along than joseph's. */
int uirk_plan_fp_line(uirk_fp_vec *out, int slice, int n, int n_dets, float axis, const float *sin_unit,
                      const float *cos_unit, const uir_subset *subs, uint32_t n_subs, int max_angles, int wmax);
/* tiles of 64 x tile_rows pixels */
int uirk_plan_bp(uirk_bp_vec *out, int slice, int n, int tile_rows);
void uirk_plan_trig(const float *sin_lut, const float *cos_lut, uint32_t n, float *trig);
/* One workgroup per item, each on the XCD whose range holds it: n_xcd times
 * the largest range (workgroups past their XCD's range exit at once), so
 * that the hardware's dispatch balances the items' uneven costs. */
int uirk_plan_grid_xcd(const uirk_xcd_part *part, int n_items);
int uirk_plan_num_xcd(void);
/* Splits items [0, n) into n_xcd contiguous ranges of about equal total cost
 * (cost[i] >= 0; NULL = unit cost). */
void uirk_plan_xcd(const double *cost, int n, int n_xcd, uirk_xcd_part *out);
double uirk_fp_item_cost(const uirk_fp_item *it);
/* BP slice pairs halve the workgroups of a launch and hold fewer waves per
 * SIMD, so they are used only while the paired table keeps at least this
 * many items: 4 workgroups per CU of the current device, or the value of the
 * environment variable UIR_PAIR_MIN_ITEMS when set (0: pair whenever two
 * slices of one geometry are adjacent). */
int uirk_plan_pair_min_items(void);
/* 1 if a BP table of this many items (at the tuned tile rows, before
 * pairing) holds no more than uirk_plan_pair_min_items(): too few workgroups
 * to fill the device, so its tiles take 32 rows instead of 64.  A pixel's BP
 * result does not depend on its tile's rows (its window and sample
 * coordinates are those of its 64 x 64 super-tile). */
int uirk_plan_bp_rows32(size_t items);
/* FP32 <-> storage conversions */
int uirk_convert(const void *src, int src_dtype, void *dst, int dst_dtype, size_t n, void *stream);

/* Split-Bregman TV with SIRT inner sweeps, per N x N slice.  u is in the
 * storage dtype; ex = dx - bx, ey = dy - by, bx, by and tv are FP32.
 * sb_tv:     tv = Dx^T (ex - Dx u) + Dy^T (ey - Dy u)
 * sb_shrink: g = D u + b, t = max(|g| - 1/lambda, 0) / max(|g|, 1e-12),
 *            d = t g, b = g - d, e = d - b
 * (the periodic differences of ufo-ir-gradient-processor.cl:1-51:
 * (Dx u)[y][x] = u[y][x] - u[y][x-1], (Dx^T v)[y][x] = v[y][x] - v[y][x+1]). */
int uirk_sb_tv(const void *u, int dtype, const float *ex, const float *ey, float *tv, uint32_t n, void *stream);
int uirk_sb_shrink(const void *u, int dtype, float *bx, float *by, float *ex, float *ey, uint32_t n, float lambda,
                   void *stream);

/* ---- Batched ASD-POCS steps between the projections (uir_asd.hip) ----
 * Grids are (nblk, n_slices): block (c, s) covers elements [c, c + 1) *
 * UIRK_ASD_CHUNK of slice s (TVGRAD: a 256 x UIRK_ASD_TV_ROWS tile; nblk
 * must cover the largest image both ways; sinograms are strided over the
 * nblk blocks) and writes partials[s * nblk + c]. */
#define UIRK_ASD_CHUNK 4096
#define UIRK_ASD_TV_ROWS 16 /* TVGRAD blocks: 256 columns x 16 rows */
typedef struct {
    void *x, *x_prev;    /* image and its checkpoint (image dtype) */
    float *g;            /* TV gradient (FP32) */
    const void *r;       /* residual sinogram b - A x (sinogram dtype) */
    uint32_t n, n_sino;  /* image edge; n_angles * n_dets */
} uirk_asd_slice;
typedef struct {
    float beta, relax, dtgv, dd, dp, dg, factor;
    int32_t it, stop, pad;
} uirk_asd_state;
typedef struct {
    float beta_red, alpha, alpha_red, r_max;
    int32_t textbook;
} uirk_asd_params;
/* kind: BEGIN (flag 0: x = 0, 1: x_prev = x, 2: x = x_prev, 3: x = g as image
 * dtype), POSC_DP (flag = posc:
 * x = max(x, 0), partial |x - x_prev|, x_prev = x), DIFF (partial |x - x_prev|),
 * TVGRAD (g = l1_grad(x), partial |g|), TVSTEP (x -= factor g); l2 selects
 * sums of squares over sums of magnitudes. */
enum { UIRK_ASD_BEGIN = 0, UIRK_ASD_POSC_DP = 1, UIRK_ASD_DIFF = 2, UIRK_ASD_TVGRAD = 3, UIRK_ASD_TVSTEP = 4 };
int uirk_asd_image(int kind, int img_dtype, const uirk_asd_slice *d_sl, int n_slices, int nblk,
                   const uirk_asd_state *d_state, int flag, int l2, double *d_partials, void *stream);
/* partial |r| over the residual sinograms */
int uirk_asd_sino_norm(int dtype, const uirk_asd_slice *d_sl, int n_slices, int nblk, int l2, double *d_partials,
                       void *stream);
/* what: DD, DP (sets dtgv = alpha dp in iteration 0), G (factor = dtgv / |g|;
 * textbook: a zero gradient ends the TV steps), DG (diagnostics row, dtgv
 * and beta updates, the next data step's relaxation, iteration count). */
enum { UIRK_ASD_DD = 0, UIRK_ASD_DP = 1, UIRK_ASD_G = 2, UIRK_ASD_DG = 3 };
int uirk_asd_finish(uirk_asd_state *d_state, const double *d_partials, int n_slices, int nblk, int what,
                    const uirk_asd_params *p, float *d_diag, int max_iters, void *stream);

/* Fused SART step (uir_fast.hip): one pass over the images applies the BP of
 * bp_angle (x_out = x_in + relax A^T r[bp_angle], relax from the slice's
 * state; bp_angle < 0 or >= n_angles: x_out = x_in) and writes the partial
 * FP sums of fp_angle (< 0: none) per (band of 64 lines, tile % 3, ray) to
 * part; uirk_sart_rows then forms r[fp_angle] = (b + corr FP) * ray_w.
 * x_in and x_out are distinct buffers (tiles read a halo of their
 * neighbours).  Tiles are 64 x 64; part holds ceil(n / 64) * 3 * n_dets
 * floats per slice. */
typedef struct {
    const void *x_in;    /* image dtype */
    void *x_out;
    void *r;             /* residual rows (sinogram dtype) */
    const void *b;       /* measured sinogram (sinogram dtype) */
    const float *ray_w;
    const float *trig;
    float *part;
    int32_t n, n_dets, n_angles, pad;
    float axis, corr;
    const uirk_line_angle *line; /* the line model (uirk_sart_line) only */
} uirk_sart_slice;
int uirk_sart_step(const uirk_sart_slice *d_slices, const uirk_asd_state *d_state, int n_slices, int max_n,
                   int dtype, int img_dtype, int bp_angle, int fp_angle, int *d_err, void *stream);
int uirk_sart_rows(const uirk_sart_slice *d_slices, int n_slices, int max_d, int dtype, int angle, void *stream);
/* The same SART angle for slices of the line model (uir_line.hip), in two
 * launches: r[angle] = (b + corr FP) * ray_w, then x_out = x_in + relax
 * A^T r[angle] pixel by pixel (x_out may equal x_in); descriptors need line
 * (the geometry's angle table) and trig (its unit directions). */
int uirk_sart_line(const uirk_sart_slice *d_slices, const uirk_asd_state *d_state, int n_slices, int max_n, int max_d,
                   int dtype, int img_dtype, int angle, void *stream);

/* Finite differences (ufo-ir-gradient-processor.cl) and l1_grad. */
int uirk_diff(int kind, const float *in, float *out, uint32_t w, uint32_t h, void *stream);
enum { UIRK_DX = 0, UIRK_DXT = 1, UIRK_DY = 2, UIRK_DYT = 3 };
int uirk_l1_grad(const float *in, float *out, uint32_t w, uint32_t h, void *stream);

#ifdef __cplusplus
}
#endif
#endif
