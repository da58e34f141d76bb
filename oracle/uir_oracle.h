/*
 * uir_oracle.h -- CPU oracle for the parallel-beam iterative reconstruction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only the
 * test suite, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
 * load it.  The product library (libuir.so) never links or calls it.
 *
 * Two oracles live here:
 *   orc_*   -- a "ref-compat" FP32 restatement of ufo-ir's OpenCL kernels and
 *              method loops, operation for operation, with the same rounding
 *              points (float incremental ray stepping, double-promoted start
 *              coordinate, 64-sample blocked sums, per-run BP accumulation,
 *              sequential float host reductions).  Every function cites the
 *              reference file:line it follows.
 *   orc64_* -- the same Joseph operator pair evaluated in FP64 with exact
 *              coordinates (no incremental rounding); the accuracy reference
 *              for reduced-precision device paths.
 *
 * Image layout: [y][x] row-major, N x N (the reference always uses
 * N == n_dets, ufo-ir-parallel-projector-task.c:339-346; N != n_dets is the
 * generalisation documented in orc_geom).  Sinogram layout: [angle][det].
 */
#ifndef UIR_ORACLE_H
#define UIR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* projector-parallel-joseph.cl:25-34 / ufo-ir-parallel-projector-task.h:37-53 */
typedef struct {
    uint32_t offset;
    uint32_t n;
    int32_t direction; /* 1 = Vertical (|sin| <= |cos|), 0 = Horizontal */
} orc_subset;

/*
 * Geometry.  n == n_dets reproduces the reference exactly (its FP sets
 * width = height = n_dets, ufo-ir-parallel-projector-task.c:575-578).  For
 * n != n_dets every place where the reference kernels use the image width as
 * a detector quantity (diff, half_active_dets, the BP detector shift) uses
 * n_dets, and every place where they use it as an image quantity (origin
 * shift, sample count, slice height) uses n.
 */
typedef struct {
    int n;            /* image edge */
    int n_dets;       /* detector bins */
    int n_angles;     /* rows of the sinogram */
    float axis;       /* resolved axis position (>= 0) */
    const float *sin_lut;
    const float *cos_lut;
    int model;        /* projection model of the FP64 operators: 0 = joseph
                       * (the reference's), 1 = line (exact ray-pixel
                       * intersection lengths; no reference counterpart) */
} orc_geom;

/* ---- geometry (ufo-ir-parallel-projector-task.c) ---- */
void  orc_lut(uint32_t n_angles, float step, float *sin_lut, float *cos_lut);
float orc_axis_resolve(float axis_property, uint32_t n_dets);
int   orc_runs(const float *sin_lut, const float *cos_lut, uint32_t n_angles, orc_subset *out);
int   orc_angle_subsets(const float *sin_lut, const float *cos_lut, uint32_t n_angles, orc_subset *out);

/* ---- Joseph projector, ref-compat FP32 (projector-parallel-joseph.cl) ---- */
void orc_fp_subset(const orc_geom *g, const float *vol, float *sino, orc_subset sub, float corr);
void orc_bp_subset(const orc_geom *g, float *vol, const float *sino, orc_subset sub, float relax);
void orc_fp(const orc_geom *g, const float *vol, float *sino, const orc_subset *runs, int n_runs, float corr);
void orc_bp(const orc_geom *g, float *vol, const float *sino, const orc_subset *runs, int n_runs, float relax);

/* ---- elementwise ops (ufo-ir-basic-ops.cl) ---- */
void orc_op_set(float *out, size_t n, float value);
void orc_op_inv(float *inout, size_t n);
void orc_op_mul(const float *a, const float *b, float *out, size_t n);
void orc_op_add(const float *a, const float *b, float *out, size_t n);
void orc_op_sub(const float *a, const float *b, float *out, size_t n);
void orc_op_sub_scaled(const float *a, const float *b, float m, float *out, size_t n);
void orc_op_add_scaled(const float *a, const float *b, float m, float *out, size_t n);
void orc_op_mul_rows(const float *a, const float *b, float *out, size_t width, uint32_t offset, uint32_t rows);
void orc_op_posc(const float *in, float *out, size_t n);

/* ---- host reductions and host elementwise (ufo-ir-basic-ops*.c) ---- */
float orc_l1_norm(const float *v, size_t w, size_t h);
float orc_dot_compat(const float *a, const float *b, size_t w, size_t h);
float orc_l2_compat(const float *a, size_t w, size_t h);
void  orc_mul_scalar(float *v, size_t n, float m);
void  orc_normalize(float *v, size_t n);
void  orc_sqrt(float *v, size_t n);
void  orc_mul_ew(const float *a, const float *b, float *out, size_t n);
void  orc_max_ew(const float *a, const float *b, float *out, size_t n);
void  orc_div_ew(const float *a, const float *b, float *out, size_t n);

/* ---- finite differences and TV (ufo-ir-gradient-processor.cl, ufo-math-tvstd-method.cl) ---- */
void orc_dx(const float *in, float *out, int w, int h);
void orc_dxt(const float *in, float *out, int w, int h);
void orc_dy(const float *in, float *out, int w, int h);
void orc_dyt(const float *in, float *out, int w, int h);
void orc_l1_grad(const float *in, float *out, int w, int h);

/* ---- methods (src/tasks/ufo-ir-*-task.c) ---- */
void orc_sirt(const orc_geom *g, const float *b, float *x, uint32_t iters, float relax, int posc);
void orc_sart(const orc_geom *g, const float *b, float *x, uint32_t iters, float relax);

typedef struct {
    float beta, beta_red, alpha, alpha_red, r_max;
    int positive_constraint;
    int df_is_sirt;          /* 0: SART df-minimizer, 1: SIRT */
    uint32_t df_iters;       /* df-minimizer num-iterations */
    float df_relax_initial;  /* df-minimizer relaxation_factor before the first set */
    uint32_t ng;             /* TV steps per iteration (textbook only; the reference takes 20) */
    int textbook;            /* 0: the reference; 1: ASD-POCS as published (see uir_oracle.c) */
} orc_asdpocs_params;
/* diag (optional, 4 floats per outer iteration): dd, dp, dg, dtgv */
void orc_asdpocs(const orc_geom *g, const float *b, float *out, uint32_t iters,
                 const orc_asdpocs_params *p, float *diag);

/* b is normalised in place, as the reference does to its input buffer.
 * dot_mode 0: the reference's strided "dot" (ufo-ir-basic-ops-processor.c:181-193);
 * dot_mode 1: a plain sequential float dot (test switch for pinning CGS). */
void orc_sbtv(const orc_geom *g, float *b, float *u, uint32_t iters, float lambda, float mu, int dot_mode);
/* one CGS solve with the SBTV operator, exposed for pinning */
int orc_sbtv_cgs(const orc_geom *g, const float *rhs, float *x, const float *x0, uint32_t max_iter,
                 float lambda, float mu, int dot_mode);
void orc_sbtv_apply(const orc_geom *g, const float *in, float *out, float lambda, float mu);
/* SIRT pixel weights of the line model: C = 1 / BP(1) where BP(1) >=
 * ORC_LINE_PIX_MIN, else 0.  A pixel that rays only touch at a corner has
 * BP(1) = 0 exactly and ~1e-7 in FP32 arithmetic; the floor keeps FP32 from
 * giving it C ~ 1e7 and a full-size update. */
#define ORC_LINE_PIX_MIN (1.0 / 1024.0)
/* The line model's direction: the LUT pair scaled to unit length, with a
 * component below 2^-20 taken as exactly 0 (and the other as +-1). */
void orc64_line_dir(float cos_lut, float sin_lut, double *c, double *s);
/* The line model's weight: the length of the line X cos - Y sin = t inside
 * the unit pixel whose centre projects to t - d. */
double orc64_line_chord(double d, double c, double s);

/* The same CGS (ufo-ir-sbtv-task.c:404-585) with a dense (w h) x (w h)
 * operator M (row-major, each row summed in float in order): for pinning the
 * CGS control flow on constructed systems.  Returns the flag. */
int orc_cgs_dense(const float *M, uint32_t w, uint32_t h, const float *b, float *x, const float *x0,
                  uint32_t max_iter, int dot_mode);
/* the shrinkage step of update_db on (tx, ty) -> (dx, dy) */
void orc_sbtv_shrink(const float *tx, const float *ty, size_t n, float lambda, float *dx, float *dy);

/* ---- FP64 evaluation of the same operators (exact coordinates) ---- */
void orc64_fp(const orc_geom *g, const double *vol, double *sino);
void orc64_bp(const orc_geom *g, double *vol, const double *sino);
/* r_min: rays with FP(1) < r_min get weight 0 (0 = only FP(1) = 0 does) */
void orc64_sirt(const orc_geom *g, const double *b, double *x, uint32_t iters, double relax, int posc,
                double r_min);
void orc64_sart(const orc_geom *g, const double *b, double *x, uint32_t iters, double relax);
void orc64_l1_grad(const double *in, double *out, int n);
/* ASD-POCS (both variants of orc_asdpocs); diag as orc_asdpocs */
void orc64_asdpocs(const orc_geom *g, const double *b, double *out, uint32_t iters, const orc_asdpocs_params *p,
                   double r_min, double *diag);

/* split-Bregman TV with SIRT-preconditioned sweeps for the u-subproblem (FP64):
 * outer Bregman iterations x inner sweeps, relaxation w, TV weight lambda,
 * data weight mu; orc64_sb_tv / orc64_sb_shrink are its two TV steps.  The
 * preconditioner is P = 1 / (mu A + 8 lambda) for joseph (A angles) and, per
 * pixel, P = 1 / (mu BP(1) + 8 lambda) for the line model, with BP(1) taken
 * as 0 below ORC_LINE_PIX_MIN (then P = 1 / (8 lambda), or 0 for lambda = 0,
 * as the line SIRT's pixel weight). */
void orc64_sb_tv(const double *u, const double *ex, const double *ey, double *tv, int n);
void orc64_sb_shrink(const double *u, double *bx, double *by, double *ex, double *ey, int n, double lambda);
void orc64_sbsirt(const orc_geom *g, const double *f, double *u, uint32_t outer, uint32_t inner, double relax,
                  double lambda, double mu, int posc, double r_min);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
