/*
 * uir_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see uir_oracle.h).
 *
 * Ref-compat FP32 restatement of ufo-ir's parallel-beam path plus an FP64
 * evaluation of the same operators.  Built with -ffp-contract=off so every
 * float operation rounds where the reference source rounds.  All reference
 * citations are relative to the ufo-ir source tree.
 */
#include "uir_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_BLOCK 64 /* projector-parallel-joseph.cl:4 */

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Geometry                                                                 */
/* ------------------------------------------------------------------------ */

/* create_lut_buffer, ufo-ir-parallel-projector-task.c:436-460:
 * rad_angle = i * angles_step is a float product (guint * gfloat) stored in a
 * double; the LUT holds (gfloat) sin/cos((double) rad_angle). */
void orc_lut(uint32_t n_angles, float step, float *sin_lut, float *cos_lut)
{
    for (uint32_t i = 0; i < n_angles; ++i) {
        double rad = (double)((float)i * step);
        sin_lut[i] = (float)sin(rad);
        cos_lut[i] = (float)cos(rad);
    }
}

/* subset_fp_real / subset_bp_real, ufo-ir-parallel-projector-task.c:524-527,
 * 584-587: a negative axis_position means n_dets / 2.0 (double) stored in a float. */
float orc_axis_resolve(float axis_property, uint32_t n_dets)
{
    if (axis_property < 0)
        return (float)(n_dets / 2.0);
    return axis_property;
}

/* generate_full_subsets_list, ufo-ir-parallel-projector-task.c:462-498 (and the
 * identical ufo-ir-asdpocs-task.c:512-548): runs of equal direction, with
 * direction = fabs(sin) <= fabs(cos) (Vertical == 1). */
int orc_runs(const float *s, const float *c, uint32_t n_angles, orc_subset *out)
{
    if (n_angles == 0)
        return 0;
    int k = 0;
    out[0].direction = fabs(s[0]) <= fabs(c[0]);
    out[0].offset = 0;
    out[0].n = 1;
    for (uint32_t i = 1; i < n_angles; ++i) {
        int d = fabs(s[i]) <= fabs(c[i]);
        if (d != out[k].direction) {
            ++k;
            out[k].direction = d;
            out[k].offset = i;
            out[k].n = 1;
        } else {
            out[k].n++;
        }
    }
    return k + 1;
}

/* generate_subsets, ufo-ir-sart-task.c:230-248: one subset per angle. */
int orc_angle_subsets(const float *s, const float *c, uint32_t n_angles, orc_subset *out)
{
    for (uint32_t i = 0; i < n_angles; ++i) {
        out[i].direction = fabs(s[i]) <= fabs(c[i]);
        out[i].offset = i;
        out[i].n = 1;
    }
    return (int)n_angles;
}

/* ------------------------------------------------------------------------ */
/* OpenCL image sampling restated (OpenCL 1.2 spec, section 8.2)            */
/* ------------------------------------------------------------------------ */

/* CLK_ADDRESS_CLAMP: texels outside the image read as border colour 0. */
static inline float texel(const float *img, int w, int h, int i, int j)
{
    return (i >= 0 && i < w && j >= 0 && j < h) ? img[(size_t)j * w + i] : 0.0f;
}

/* CLK_FILTER_LINEAR, unnormalised coordinates: i0 = floor(u - 0.5),
 * a = frac(u - 0.5), weighted sum (1-a)(1-b)T00 + a(1-b)T10 + (1-a)bT01 + abT11.
 * Used by FP_hor/FP_vert (projector-parallel-joseph.cl:85,146) and BP (:212). */
static inline float read_linear(const float *img, int w, int h, float x, float y)
{
    float u = x - 0.5f, v = y - 0.5f;
    float fu = floorf(u), fv = floorf(v);
    int i0 = (int)fu, j0 = (int)fv;
    float a = u - fu, b = v - fv;
    float oma = 1.0f - a, omb = 1.0f - b;
    float t00 = texel(img, w, h, i0, j0), t10 = texel(img, w, h, i0 + 1, j0);
    float t01 = texel(img, w, h, i0, j0 + 1), t11 = texel(img, w, h, i0 + 1, j0 + 1);
    return ((oma * omb * t00 + a * omb * t10) + oma * b * t01) + a * b * t11;
}

/* ------------------------------------------------------------------------ */
/* Joseph forward projection (projector-parallel-joseph.cl:36-156)          */
/* ------------------------------------------------------------------------ */

/* One subset launch of FP_hor (direction 0) or FP_vert (direction 1), global
 * size (n_dets, subset.n) as bound by subset_fp_real
 * (ufo-ir-parallel-projector-task.c:562-625).  sino is read (r_sinogram) and
 * written (w_sinogram) in place: w = r + corr * ray_sum. */
void orc_fp_subset(const orc_geom *g, const float *vol, float *sino, orc_subset sub, float corr)
{
    const int N = g->n, D = g->n_dets;
    const float axis = g->axis;
    const float required_width = axis * 2;           /* :51 */
    const float diff = (float)D - required_width;     /* :55 (reference width == n_dets) */
    const float rot = diff / 2.0f;                    /* :58 */
    const float origin_shift = (float)N / 2.0f;       /* :60 */
    const int n_blocks = (int)ceilf((float)N / ORC_BLOCK); /* :81 */

#pragma omp parallel for collapse(2) schedule(static)
    for (uint32_t k = 0; k < sub.n; ++k) {
        for (int s = 0; s < D; ++s) {
            const int ang = (int)(sub.offset + k);
            /* Switch off inactive detectors, :63-66 */
            if ((diff < 0 && (float)s < fabsf(diff)) || (diff > 0 && (float)s > required_width))
                continue;
            const float st = g->sin_lut[ang], ct = g->cos_lut[ang];
            /* The start coordinate is a double expression (unsuffixed 0.5 literal,
             * :73 / :134) rounded to float on assignment. */
            const double t = ((0.5 + (double)rot) + (double)s) - (double)(0.5f * (float)D);
            float x, y, step;
            if (sub.direction == 0) { /* FP_hor, :68-75 */
                const float det_step = -1.0f / st;
                step = ct / st;
                x = 0.5f;
                y = (float)(t * (double)det_step + (double)((-origin_shift) * step)
                            + (double)(0.5f * (float)N));
            } else { /* FP_vert, :129-136 */
                const float det_step = 1.0f / ct;
                step = st / ct;
                y = 0.5f;
                x = (float)(t * (double)det_step + (double)((-origin_shift) * step)
                            + (double)(0.5f * (float)N));
            }
            float detected = 0.0f;
            for (int j = 0; j < n_blocks; ++j) { /* :82-90 */
                float inner = 0.0f;
                if (sub.direction == 0) {
                    for (int i = 0; i < ORC_BLOCK; ++i) {
                        inner += read_linear(vol, N, N, x, y);
                        y += step;
                        x += 1.0f;
                    }
                } else {
                    for (int i = 0; i < ORC_BLOCK; ++i) {
                        inner += read_linear(vol, N, N, x, y);
                        x += step;
                        y += 1.0f;
                    }
                }
                detected += inner;
            }
            float *out = &sino[(size_t)ang * D + s];
            *out = *out + detected * corr; /* :92-94 */
        }
    }
}

/* One subset launch of BP (projector-parallel-joseph.cl:158-218), global size
 * = volume dims (subset_bp_real, ufo-ir-parallel-projector-task.c:500-560).
 * vol is read (r_volume) and written (w_volume) in place: w = r + relax * sum. */
void orc_bp_subset(const orc_geom *g, float *vol, const float *sino, orc_subset sub, float relax)
{
    const int N = g->n, D = g->n_dets, A = g->n_angles;
    const float axis = g->axis;
    const float required_width = axis * 2;                        /* :174 */
    const float diff = (float)D - required_width;                  /* :178 */
    const float half_active = diff < 0 ? (float)D - axis : axis;   /* :180 */
    const float edge0 = axis - half_active + 0.5f;                 /* :181 */
    const float edge1 = axis + half_active - 0.5f;                 /* :182 */
    const float rot = diff / 2.0f;                                 /* :185 */
    const float det_shift = (float)D / 2.0f;                       /* :187 as a detector quantity */
    const float origin_shift = (float)N / 2.0f;                    /* :187 as an image quantity */
    const float shift = det_shift - rot;                           /* :201 */

#pragma omp parallel for collapse(2) schedule(static)
    for (int py = 0; py < N; ++py) {
        for (int px = 0; px < N; ++px) {
            const float fX = (float)px + 0.5f - origin_shift;      /* :189 */
            const float fY = (float)py + 0.5f - origin_shift;      /* :190 */
            float value = 0.0f;
            float sy = (float)sub.offset + 0.5f;                   /* :194 */
            for (uint32_t i = 0; i < sub.n; ++i) {                 /* :196-214 */
                const float st = g->sin_lut[i + sub.offset];
                const float ct = g->cos_lut[i + sub.offset];
                float u = fX * ct - fY * st + shift;
                if (u > edge1)
                    u = edge1;
                else if (u < edge0)
                    u = edge0;
                value += read_linear(sino, D, A, u, sy);
                sy += 1.0f;
            }
            float *out = &vol[(size_t)py * N + px];
            *out = *out + relax * value;                           /* :216-217 */
        }
    }
}

/* forward / backward, ufo-ir-parallel-projector-task.c:390-429: one launch per run. */
void orc_fp(const orc_geom *g, const float *vol, float *sino, const orc_subset *runs, int n_runs, float corr)
{
    for (int i = 0; i < n_runs; ++i)
        orc_fp_subset(g, vol, sino, runs[i], corr);
}

void orc_bp(const orc_geom *g, float *vol, const float *sino, const orc_subset *runs, int n_runs, float relax)
{
    for (int i = 0; i < n_runs; ++i)
        orc_bp_subset(g, vol, sino, runs[i], relax);
}

/* ------------------------------------------------------------------------ */
/* Elementwise kernels, ufo-ir-basic-ops.cl                                  */
/* ------------------------------------------------------------------------ */

void orc_op_set(float *out, size_t n, float value) /* operation_set :4-16 */
{
    for (size_t i = 0; i < n; ++i)
        out[i] = value;
}

void orc_op_inv(float *v, size_t n) /* operation_inv :18-36 (in place, ufo-ir-basic-ops.c:158-159) */
{
    for (size_t i = 0; i < n; ++i)
        v[i] = (v[i] != 0) ? 1.0f / v[i] : 0;
}

void orc_op_mul(const float *a, const float *b, float *out, size_t n) /* operation_mul :38-58 */
{
    for (size_t i = 0; i < n; ++i)
        out[i] = a[i] * b[i];
}

void orc_op_add(const float *a, const float *b, float *out, size_t n) /* operation_add :60-80 */
{
    for (size_t i = 0; i < n; ++i)
        out[i] = a[i] + b[i];
}

void orc_op_sub(const float *a, const float *b, float *out, size_t n) /* operation_deduction :82-102 */
{
    for (size_t i = 0; i < n; ++i)
        out[i] = a[i] - b[i];
}

void orc_op_sub_scaled(const float *a, const float *b, float m, float *out, size_t n) /* operation_deduction2 :104-125 */
{
    for (size_t i = 0; i < n; ++i)
        out[i] = a[i] - m * b[i];
}

void orc_op_add_scaled(const float *a, const float *b, float m, float *out, size_t n) /* operation_add2 :127-148 */
{
    for (size_t i = 0; i < n; ++i)
        out[i] = a[i] + m * b[i];
}

/* op_mulRows :150-171, global size (width, rows) at row offset. */
void orc_op_mul_rows(const float *a, const float *b, float *out, size_t width, uint32_t offset, uint32_t rows)
{
    for (size_t y = offset; y < (size_t)offset + rows; ++y)
        for (size_t x = 0; x < width; ++x)
            out[y * width + x] = a[y * width + x] * b[y * width + x];
}

void orc_op_posc(const float *in, float *out, size_t n) /* POSC :255-273 */
{
    for (size_t i = 0; i < n; ++i)
        out[i] = in[i] > 0 ? in[i] : 0;
}

/* ------------------------------------------------------------------------ */
/* Host-side reductions and elementwise loops                               */
/* ------------------------------------------------------------------------ */

/* ufo_ir_op_l1_norm, ufo-ir-basic-ops.c:259-277: sequential float sum with
 * index i * dims[1] + j over i < dims[0], j < dims[1] (linear order). */
float orc_l1_norm(const float *v, size_t w, size_t h)
{
    float norm = 0;
    for (size_t i = 0; i < w; ++i)
        for (size_t j = 0; j < h; ++j)
            norm += (float)fabs(v[i * h + j]);
    return norm;
}

/* ufo_ir_basic_ops_processor_dot_product, ufo-ir-basic-ops-processor.c:148-196:
 * partsCnt = dims[1], partLen = len / partsCnt, element index partNum * i. */
float orc_dot_compat(const float *a, const float *b, size_t w, size_t h)
{
    const size_t parts = h, part_len = (w * h) / h;
    float norm = 0;
    for (size_t p = 0; p < parts; ++p) {
        float part = 0;
        for (size_t i = 0; i < part_len; ++i) {
            size_t idx = p * i;
            part += a[idx] * b[idx];
        }
        norm += part;
    }
    return norm;
}

/* ufo_ir_basic_ops_processor_l2_norm, :239-247: sqrt (double) of the "dot". */
float orc_l2_compat(const float *a, size_t w, size_t h)
{
    return (float)sqrt(orc_dot_compat(a, a, w, h));
}

/* The inner product of the textbook variants (compat = false): every element
 * once, accumulated in double and rounded to float once. */
static float dot_plain(const float *a, const float *b, size_t n)
{
    double s = 0;
    for (size_t i = 0; i < n; ++i)
        s += (double)a[i] * (double)b[i];
    return (float)s;
}

static float l2_plain(const float *a, size_t n) { return (float)sqrt((double)dot_plain(a, a, n)); }

void orc_mul_scalar(float *v, size_t n, float m) /* :324-340 */
{
    for (size_t i = 0; i < n; ++i)
        v[i] *= m;
}

/* ufo_ir_basic_ops_processor_normalization, :342-367: min-max in place. */
void orc_normalize(float *v, size_t n)
{
    float mx = -3.40282347e+38f, mn = 3.40282347e+38f;
    for (size_t i = 0; i < n; ++i) {
        mx = (float)fmax(v[i], mx);
        mn = (float)fmin(v[i], mn);
    }
    float delta = 1 / (mx - mn);
    for (size_t i = 0; i < n; ++i)
        v[i] = delta * v[i] - mn * delta;
}

void orc_sqrt(float *v, size_t n) /* :415-430 */
{
    for (size_t i = 0; i < n; ++i)
        v[i] = (float)sqrt(v[i]);
}

void orc_mul_ew(const float *a, const float *b, float *out, size_t n) /* :583-588 */
{
    for (size_t i = 0; i < n; ++i)
        out[i] = a[i] * b[i];
}

void orc_max_ew(const float *a, const float *b, float *out, size_t n) /* :590-596 */
{
    for (size_t i = 0; i < n; ++i)
        out[i] = (float)fmax(a[i], b[i]);
}

void orc_div_ew(const float *a, const float *b, float *out, size_t n) /* :598-604 */
{
    for (size_t i = 0; i < n; ++i)
        out[i] = a[i] / b[i];
}

/* ------------------------------------------------------------------------ */
/* Finite differences (ufo-ir-gradient-processor.cl) and TV gradient         */
/* ------------------------------------------------------------------------ */

void orc_dx(const float *in, float *out, int w, int h) /* Dx :1-11 */
{
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int idx = y * w + x;
            out[idx] = in[idx] - in[idx - 1 + (x == 0 ? w : 0)];
        }
}

void orc_dxt(const float *in, float *out, int w, int h) /* Dxt :13-24, stopInd = w-1 */
{
    const int stop = w - 1;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int idx = y * w + x;
            out[idx] = in[idx] - in[idx + 1 - (x == stop ? w : 0)];
        }
}

void orc_dy(const float *in, float *out, int w, int h) /* Dy :26-37, lastOffset = w*h */
{
    const int last = w * h;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int idx = y * w + x;
            out[idx] = in[idx] - in[idx - w + (y == 0 ? last : 0)];
        }
}

void orc_dyt(const float *in, float *out, int w, int h) /* Dyt :39-51, stopInd = h-1 */
{
    const int last = w * h, stop = h - 1;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int idx = y * w + x;
            out[idx] = in[idx] - in[idx + w - (y == stop ? last : 0)];
        }
}

static inline float nb_read(const float *img, int w, int h, int x, int y)
{
    return (x >= 0 && x < w && y >= 0 && y < h) ? img[(size_t)y * w + x] : 0.0f;
}

/* l1_grad, ufo-math-tvstd-method.cl:5-57 (EPS 0.001f, nearest sampler, border 0). */
void orc_l1_grad(const float *in, float *out, int w, int h)
{
    const float eps = 0.001f;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float f3 = nb_read(in, w, h, x, y);
            float f1 = nb_read(in, w, h, x, y - 1);
            float f0 = nb_read(in, w, h, x - 1, y - 1);
            float f2 = nb_read(in, w, h, x - 1, y);
            float f4 = nb_read(in, w, h, x + 1, y);
            float f6 = nb_read(in, w, h, x + 1, y + 1);
            float f5 = nb_read(in, w, h, x, y + 1);
            float df = (2.0f * f3 - f2 - f5) / sqrtf(eps + (f3 - f2) * (f3 - f2) + (f3 - f5) * (f3 - f5))
                     - (f4 - f3) / sqrtf(eps + (f4 - f3) * (f4 - f3) + (f4 - f6) * (f4 - f6))
                     - (f1 - f3) / sqrtf(eps + (f1 - f3) * (f1 - f3) + (f1 - f0) * (f1 - f0));
            out[(size_t)y * w + x] = df;
        }
}

/* ------------------------------------------------------------------------ */
/* Methods                                                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    orc_subset *runs;
    int n_runs;
} run_list;

static run_list make_runs(const orc_geom *g)
{
    run_list r;
    r.runs = (orc_subset *)malloc(sizeof(orc_subset) * (size_t)(g->n_angles > 0 ? g->n_angles : 1));
    r.n_runs = orc_runs(g->sin_lut, g->cos_lut, (uint32_t)g->n_angles, r.runs);
    return r;
}

/* SIRT, ufo-ir-sirt-task.c:163-222.  The reference has no positivity clamp;
 * posc != 0 composes POSC (ufo-ir-basic-ops.cl:255-273) after each update. */
static void sirt_core(const orc_geom *g, const float *b, float *x, uint32_t iters, float relax, int posc,
                      int warm)
{
    const size_t nv = (size_t)g->n * g->n, ns = (size_t)g->n_angles * g->n_dets;
    run_list rl = make_runs(g);
    float *vol_tmp = (float *)malloc(nv * sizeof(float));
    float *ray_w = (float *)malloc(ns * sizeof(float));
    float *sino_tmp = (float *)malloc(ns * sizeof(float));
    float *pix_w = (float *)malloc(nv * sizeof(float));

    orc_op_set(vol_tmp, nv, 1.0f);                      /* :180-185 */
    orc_op_set(ray_w, ns, 0.0f);
    orc_fp(g, vol_tmp, ray_w, rl.runs, rl.n_runs, 1.0f);
    orc_op_inv(ray_w, ns);
    orc_op_set(sino_tmp, ns, 1.0f);                     /* :188-193 */
    orc_op_set(pix_w, nv, 0.0f);
    orc_bp(g, pix_w, sino_tmp, rl.runs, rl.n_runs, 1.0f);
    orc_op_inv(pix_w, nv);
    if (!warm)
        orc_op_set(x, nv, 0.0f);                        /* :195-197 */

    for (uint32_t it = 0; it < iters; ++it) {           /* :202-215 */
        memcpy(sino_tmp, b, ns * sizeof(float));
        orc_fp(g, x, sino_tmp, rl.runs, rl.n_runs, -1.0f);
        orc_op_mul(sino_tmp, ray_w, sino_tmp, ns);
        orc_op_set(vol_tmp, nv, 0.0f);
        orc_bp(g, vol_tmp, sino_tmp, rl.runs, rl.n_runs, relax);
        orc_op_mul(vol_tmp, pix_w, vol_tmp, nv);
        orc_op_add(vol_tmp, x, x, nv);
        if (posc)
            orc_op_posc(x, x, nv);
    }
    free(vol_tmp); free(ray_w); free(sino_tmp); free(pix_w); free(rl.runs);
}

void orc_sirt(const orc_geom *g, const float *b, float *x, uint32_t iters, float relax, int posc)
{
    sirt_core(g, b, x, iters, relax, posc, 0);
}

/* SART, ufo-ir-sart-task.c:173-228: per-angle ordered subsets.  warm != 0
 * starts from x as given instead of zero (the textbook ASD-POCS data step). */
static void sart_core(const orc_geom *g, const float *b, float *x, uint32_t iters, float relax, int warm)
{
    const size_t nv = (size_t)g->n * g->n, ns = (size_t)g->n_angles * g->n_dets;
    orc_subset *subs = (orc_subset *)malloc(sizeof(orc_subset) * (size_t)(g->n_angles > 0 ? g->n_angles : 1));
    int n_sub = orc_angle_subsets(g->sin_lut, g->cos_lut, (uint32_t)g->n_angles, subs);
    float *sino_tmp = (float *)malloc(ns * sizeof(float));
    float *vol_tmp = (float *)malloc(nv * sizeof(float));
    float *ray_w = (float *)malloc(ns * sizeof(float));

    orc_op_set(vol_tmp, nv, 1.0f);                      /* :193-200 */
    orc_op_set(ray_w, ns, 0.0f);
    for (int i = 0; i < n_sub; ++i)
        orc_fp_subset(g, vol_tmp, ray_w, subs[i], 1.0f);
    orc_op_inv(ray_w, ns);
    if (!warm)
        orc_op_set(x, nv, 0.0f);                        /* :205-207 */
    for (uint32_t it = 0; it < iters; ++it) {           /* :208-220 */
        memcpy(sino_tmp, b, ns * sizeof(float));
        for (int i = 0; i < n_sub; ++i) {
            orc_fp_subset(g, x, sino_tmp, subs[i], -1.0f);
            orc_op_mul_rows(sino_tmp, ray_w, sino_tmp, (size_t)g->n_dets, subs[i].offset, subs[i].n);
            orc_bp_subset(g, x, sino_tmp, subs[i], relax);
        }
    }
    free(subs); free(sino_tmp); free(vol_tmp); free(ray_w);
}

void orc_sart(const orc_geom *g, const float *b, float *x, uint32_t iters, float relax)
{
    sart_core(g, b, x, iters, relax, 0);
}

/* ASD-POCS, ufo-ir-asdpocs-task.c:411-509 with tvstd :550-580 (textbook = 0).
 *
 * textbook != 0 is ASD-POCS as published (Sidky and Pan, Phys. Med. Biol. 53
 * (2008) 4777, Sec. 4): the data step continues from the current image
 * instead of zero, the distances dd, dp, dg are L2 norms, dp is measured from
 * the image at the start of the iteration, each of the ng TV steps moves
 * dtgv along the L2-normalised gradient (a zero gradient ends the steps), and
 * the output is the image after the TV steps.  The gradient is the
 * reference's smoothed TV gradient (ufo-math-tvstd-method.cl:5-57). */
static float l1_or_l2(const float *v, size_t w, size_t h, int textbook)
{
    return textbook ? l2_plain(v, w * h) : orc_l1_norm(v, w, h);
}

void orc_asdpocs(const orc_geom *g, const float *b, float *out, uint32_t iters,
                 const orc_asdpocs_params *p, float *diag)
{
    const size_t nv = (size_t)g->n * g->n, ns = (size_t)g->n_angles * g->n_dets;
    const int tb = p->textbook;
    run_list rl = make_runs(g);
    float *x = (float *)calloc(nv, sizeof(float));          /* :441-442 */
    float *x_prev = (float *)calloc(nv, sizeof(float));     /* :444-445 */
    float *x_res = (float *)malloc(nv * sizeof(float));
    float *b_res = (float *)malloc(ns * sizeof(float));
    float *grad = (float *)malloc(nv * sizeof(float));
    float dp = 1.0f, dd = 1.0f, dg = 1.0f, dtgv = 1.0f;     /* :439 */
    float beta = p->beta;                                   /* :453 */
    float df_relax = p->df_relax_initial;
    const uint32_t ng = tb ? p->ng : 20;                    /* tvstd :573: 20 steps, ng unused */

    for (uint32_t it = 0; it < iters; ++it) {
        /* :459 g_object_set on a [0,1] float property: out-of-range values are
         * rejected by GLib and the previous value stays. */
        if (beta >= 0.0f && beta <= 1.0f)
            df_relax = beta;
        if (tb)
            memcpy(x_prev, x, nv * sizeof(float));
        if (p->df_is_sirt)                                  /* :460 */
            sirt_core(g, b, x, p->df_iters, df_relax, 0, tb);
        else
            sart_core(g, b, x, p->df_iters, df_relax, tb);
        if (p->positive_constraint)                         /* :463-465 */
            orc_op_posc(x, x, nv);
        if (!tb)
            memcpy(out, x, nv * sizeof(float));             /* :468 */
        memcpy(b_res, b, ns * sizeof(float));               /* :471-475 */
        orc_fp(g, x, b_res, rl.runs, rl.n_runs, -1.0f);
        dd = l1_or_l2(b_res, (size_t)g->n_dets, (size_t)g->n_angles, tb); /* :478 */
        orc_op_sub(x, x_prev, x_res, nv);                   /* :481-482 */
        dp = l1_or_l2(x_res, (size_t)g->n, (size_t)g->n, tb);
        if (it == 0)                                        /* :485-487 */
            dtgv = p->alpha * dp;
        memcpy(x_prev, x, nv * sizeof(float));              /* :490 */
        orc_op_set(grad, nv, 0.0f);                         /* tvstd :559-579 */
        for (uint32_t k = 0; k < ng; ++k) {
            orc_l1_grad(x, grad, g->n, g->n);
            float norm = l1_or_l2(grad, (size_t)g->n, (size_t)g->n, tb);
            if (tb && !(norm > 0.0f))
                break;
            float factor = dtgv / norm;
            orc_op_sub_scaled(x, grad, factor, x, nv);
        }
        orc_op_sub(x, x_prev, x_res, nv);                   /* :496-497 */
        dg = l1_or_l2(x_res, (size_t)g->n, (size_t)g->n, tb);
        beta *= p->beta_red;                                /* :498 */
        if (diag) {
            diag[4 * it + 0] = dd;
            diag[4 * it + 1] = dp;
            diag[4 * it + 2] = dg;
            diag[4 * it + 3] = dtgv;
        }
        const float epsilon = 0.001f;                       /* :495,501-503 */
        if (dg > p->r_max * dp && dd > epsilon)
            dtgv *= p->alpha_red;
        if (tb)
            memcpy(out, x, nv * sizeof(float));
    }
    free(x); free(x_prev); free(x_res); free(b_res); free(grad); free(rl.runs);
}

/* ---- SBTV (ufo-ir-sbtv-task.c) ---- */

typedef struct {
    const orc_geom *g;
    run_list rl;
    float lambda, mu;
    int dot_mode;
    size_t nv, ns;
    float *tempA, *tempAt, *tempD, *tempDxt;
} sbtv_ctx;

/* The CGS of ufo-ir-sbtv-task.c:404-585 on a vector of n = w * h floats
 * with an operator given as a callback: the SBTV system (processA) or, for
 * the oracle's own pins, a dense matrix.  dot_mode 0 is the reference's
 * strided "dot" over a (w, h) buffer (ufo-ir-basic-ops-processor.c:148-196),
 * 1 a true inner product. */
typedef void (*cgs_apply_fn)(void *ctx, const float *in, float *out);

typedef struct {
    size_t n, w, h;
    int dot_mode;
    cgs_apply_fn apply;
    void *ctx;
} cgs_op;

static float sb_dot(const cgs_op *c, const float *a, const float *b)
{
    if (c->dot_mode == 0)
        return orc_dot_compat(a, b, c->w, c->h);
    return dot_plain(a, b, c->n);
}

static float sb_l2(const cgs_op *c, const float *a)
{
    return (float)sqrt(sb_dot(c, a, a));
}

/* processA, :587-635: out = mu*At(A(in)) + lambda*(Dyt(Dy(in)) + Dxt(Dx(in))). */
static void sb_process_a(sbtv_ctx *c, const float *in, float *out)
{
    const orc_geom *g = c->g;
    orc_op_set(out, c->nv, 0.0f);
    orc_op_set(c->tempA, c->ns, 0.0f);
    orc_op_set(c->tempAt, c->nv, 0.0f);
    orc_fp(g, in, c->tempA, c->rl.runs, c->rl.n_runs, 1.0f);
    orc_bp(g, c->tempAt, c->tempA, c->rl.runs, c->rl.n_runs, 1.0f);
    orc_mul_scalar(c->tempAt, c->nv, c->mu);
    orc_dy(in, c->tempD, g->n, g->n);
    orc_dyt(c->tempD, out, g->n, g->n);
    orc_dx(in, c->tempD, g->n, g->n);
    orc_dxt(c->tempD, c->tempDxt, g->n, g->n);
    orc_op_add(out, c->tempDxt, out, c->nv);
    orc_mul_scalar(out, c->nv, c->lambda);
    orc_op_add(out, c->tempAt, out, c->nv);
}

static void sb_apply(void *ctx, const float *in, float *out) { sb_process_a((sbtv_ctx *)ctx, in, out); }

/* cgs, :404-585 (a port of MATLAB cgs with its stagnation / moresteps logic). */
static int cgs_core(const cgs_op *c, const float *b, float *x, const float *x0, uint32_t max_iter)
{
    const size_t nv = c->n;
    const float tol = 1E-06f;
    const double eps = 2.2204E-16;                          /* :35 */
    float n2b = sb_l2(c, b);
    memcpy(x, x0, nv * sizeof(float));
    int flag = 1;
    float *xmin = (float *)malloc(nv * sizeof(float));
    memcpy(xmin, x, nv * sizeof(float));
    float tolb = tol * n2b;
    float *r = (float *)malloc(nv * sizeof(float));
    c->apply(c->ctx, x, r);
    orc_op_sub(b, r, r, nv);
    float normr = sb_l2(c, r);
    float normr_act = normr;
    if (normr <= tolb) {                                    /* :435-440 */
        free(xmin); free(r);
        return 0;
    }
    float *rt = (float *)malloc(nv * sizeof(float));
    memcpy(rt, r, nv * sizeof(float));
    float normmin = normr;
    float rho = 1.0f;
    unsigned stag = 0, moresteps = 0;
    const unsigned maxmsteps = 5, maxstagsteps = 3;
    float *u = (float *)calloc(nv, sizeof(float));
    float *tmpa = (float *)calloc(nv, sizeof(float));
    float *pp = (float *)calloc(nv, sizeof(float));
    float *ph = (float *)malloc(nv * sizeof(float));
    float *q = (float *)calloc(nv, sizeof(float));
    float *vh = (float *)malloc(nv * sizeof(float));
    float *temp_sum = (float *)calloc(nv, sizeof(float));
    float *uh = (float *)malloc(nv * sizeof(float));
    float *qh = (float *)malloc(nv * sizeof(float));

    for (uint32_t it = 0; it < max_iter; ++it) {
        float rho1 = rho;
        rho = sb_dot(c, rt, r);
        if (rho == 0 || isinf(rho) || isnan(rho)) {
            flag = 4;
            break;
        }
        if (it == 0) {
            memcpy(u, r, nv * sizeof(float));
            memcpy(pp, u, nv * sizeof(float));
        } else {
            float beta = rho / rho1;
            if (beta == 0 || isinf(beta) || isnan(beta)) {
                flag = 4;
                break;
            }
            orc_op_add_scaled(r, q, beta, u, nv);
            orc_op_add_scaled(q, pp, beta, temp_sum, nv);
            orc_op_add_scaled(u, temp_sum, beta, pp, nv);
        }
        memcpy(ph, pp, nv * sizeof(float));
        c->apply(c->ctx, ph, vh);
        float rtvh = sb_dot(c, rt, vh);
        float alpha = 0;
        if (rtvh == 0) {
            flag = 4;
            break;
        } else {
            alpha = rho / rtvh;
        }
        if (isinf(alpha) || isnan(alpha)) {
            flag = 4;
            break;
        }
        orc_op_add_scaled(u, vh, -alpha, q, nv);
        orc_op_add(u, q, uh, nv);
        if (fabs(alpha) * sb_l2(c, uh) < eps * sb_l2(c, x))  /* :521 */
            stag += 1;
        else
            stag = 0;
        orc_op_add_scaled(x, uh, alpha, x, nv);
        c->apply(c->ctx, uh, qh);
        orc_op_add_scaled(r, qh, -alpha, r, nv);
        normr = sb_l2(c, r);
        normr_act = normr;
        if (normr <= tolb || stag >= maxstagsteps || moresteps) {
            c->apply(c->ctx, r, tmpa);                        /* :535, applies A to r */
            orc_op_sub(b, tmpa, r, nv);
            normr_act = sb_l2(c, r);
            if (normr_act <= tolb) {
                flag = 0;
                break;
            } else {
                if (stag >= maxstagsteps && moresteps == 0)
                    stag = 0;
                moresteps += 1;
                if (moresteps >= maxmsteps) {
                    flag = 3;
                    break;
                }
            }
        }
        if (normr_act < normmin) {
            normmin = normr_act;
            memcpy(xmin, x, nv * sizeof(float));
        }
        if (stag >= maxstagsteps) {
            flag = 3;
            break;
        }
    }
    if (flag != 0) {                                         /* :566-572 */
        c->apply(c->ctx, xmin, tmpa);
        orc_op_sub(b, tmpa, r, nv);
        if (sb_l2(c, r) <= normr_act)
            memcpy(x, xmin, nv * sizeof(float));
    }
    free(xmin); free(r); free(rt); free(u); free(tmpa); free(pp); free(ph);
    free(q); free(vh); free(temp_sum); free(uh); free(qh);
    return flag;
}

static int sb_cgs(sbtv_ctx *c, const float *b, float *x, const float *x0, uint32_t max_iter)
{
    const cgs_op op = {c->nv, (size_t)c->g->n, (size_t)c->g->n, c->dot_mode, sb_apply, c};
    return cgs_core(&op, b, x, x0, max_iter);
}

/* A dense n x n operator for the CGS pins: out_i = sum_j M_ij in_j, summed in
 * float in order of j. */
typedef struct {
    const float *M;
    size_t n;
} dense_op;

static void dense_apply(void *ctx, const float *in, float *out)
{
    const dense_op *d = (const dense_op *)ctx;
    for (size_t i = 0; i < d->n; ++i) {
        float acc = 0.0f;
        for (size_t j = 0; j < d->n; ++j)
            acc = acc + d->M[i * d->n + j] * in[j];
        out[i] = acc;
    }
}

int orc_cgs_dense(const float *M, uint32_t w, uint32_t h, const float *b, float *x, const float *x0,
                  uint32_t max_iter, int dot_mode)
{
    const size_t n = (size_t)w * h;
    dense_op d = {M, n};
    const cgs_op op = {n, w, h, dot_mode, dense_apply, &d};
    return cgs_core(&op, b, x, x0, max_iter);
}

static sbtv_ctx sb_make(const orc_geom *g, float lambda, float mu, int dot_mode)
{
    sbtv_ctx c;
    c.g = g;
    c.rl = make_runs(g);
    c.lambda = lambda;
    c.mu = mu;
    c.dot_mode = dot_mode;
    c.nv = (size_t)g->n * g->n;
    c.ns = (size_t)g->n_angles * g->n_dets;
    c.tempA = (float *)malloc(c.ns * sizeof(float));
    c.tempAt = (float *)malloc(c.nv * sizeof(float));
    c.tempD = (float *)malloc(c.nv * sizeof(float));
    c.tempDxt = (float *)malloc(c.nv * sizeof(float));
    return c;
}

static void sb_free(sbtv_ctx *c)
{
    free(c->rl.runs); free(c->tempA); free(c->tempAt); free(c->tempD); free(c->tempDxt);
}

void orc_sbtv_apply(const orc_geom *g, const float *in, float *out, float lambda, float mu)
{
    sbtv_ctx c = sb_make(g, lambda, mu, 0);
    sb_process_a(&c, in, out);
    sb_free(&c);
}

int orc_sbtv_cgs(const orc_geom *g, const float *rhs, float *x, const float *x0, uint32_t max_iter,
                 float lambda, float mu, int dot_mode)
{
    sbtv_ctx c = sb_make(g, lambda, mu, dot_mode);
    int flag = sb_cgs(&c, rhs, x, x0, max_iter);
    sb_free(&c);
    return flag;
}

/* The shrinkage of update_db, ufo-ir-sbtv-task.c:367-391, on t = (tx, ty):
 * s = sqrt(tx^2 + ty^2), k = max(s - 1/lambda, 0) / max(s, 1e-12),
 * d = k t, returned in (dx, dy). */
void orc_sbtv_shrink(const float *tx, const float *ty, size_t n, float lambda, float *dx, float *dy)
{
    const float d_lambda = -1 / lambda;
    for (size_t i = 0; i < n; ++i) {
        float s = tx[i] * tx[i];
        s = s + ty[i] * ty[i];
        s = (float)sqrt(s);
        float top = d_lambda + s;
        top = (float)fmax(top, 0.0f);
        const float den = (float)fmax(s, 1E-12f);
        const float k = top / den;
        dx[i] = k * tx[i];
        dy[i] = k * ty[i];
    }
}

/* SBTV process, ufo-ir-sbtv-task.c:230-302 with calculate_b :304-334 and
 * update_db :336-402. */
void orc_sbtv(const orc_geom *g, float *input, float *u, uint32_t iters, float lambda, float mu, int dot_mode)
{
    sbtv_ctx c = sb_make(g, lambda, mu, dot_mode);
    const size_t nv = c.nv, ns = c.ns;
    const int N = g->n;
    orc_normalize(input, ns);                               /* :239 */
    float *f = (float *)malloc(ns * sizeof(float));
    memcpy(f, input, ns * sizeof(float));                   /* :246-248 */
    orc_normalize(f, ns);
    float *fbp = (float *)calloc(nv, sizeof(float));        /* :251-253 */
    orc_bp(g, fbp, f, c.rl.runs, c.rl.n_runs, 1.0f);
    orc_op_set(u, nv, 0.0f);                                /* :256-257 */
    float *up = (float *)malloc(nv * sizeof(float));
    float *bb = (float *)malloc(nv * sizeof(float));
    float *bx = (float *)calloc(nv, sizeof(float));
    float *by = (float *)calloc(nv, sizeof(float));
    float *dx = (float *)calloc(nv, sizeof(float));
    float *dy = (float *)calloc(nv, sizeof(float));
    float *tmpx = (float *)malloc(nv * sizeof(float));
    float *tmpy = (float *)malloc(nv * sizeof(float));
    float *tmp = (float *)malloc(nv * sizeof(float));
    orc_mul_scalar(fbp, nv, mu);                            /* :279 */

    for (uint32_t it = 0; it < iters; ++it) {               /* :282-289 */
        memcpy(up, u, nv * sizeof(float));
        /* calculate_b */
        orc_op_sub(dx, bx, tmp, nv);
        orc_dxt(tmp, tmpx, N, N);
        orc_op_sub(dy, by, tmp, nv);
        orc_dyt(tmp, tmpy, N, N);
        orc_op_add(tmpx, tmpy, bb, nv);
        orc_mul_scalar(bb, nv, lambda);
        orc_op_add(fbp, bb, bb, nv);
        /* cgs */
        sb_cgs(&c, bb, u, up, 30);
        /* update_db */
        orc_dx(u, tmpx, N, N);
        orc_op_add(tmpx, bx, tmpx, nv);
        orc_dy(u, tmpy, N, N);
        orc_op_add(tmpy, by, tmpy, nv);
        orc_sbtv_shrink(tmpx, tmpy, nv, lambda, dx, dy);     /* :367-385 */
        orc_op_sub(tmpx, dx, bx, nv);
        orc_op_sub(tmpy, dy, by, nv);
    }
    free(f); free(fbp); free(up); free(bb); free(bx); free(by); free(dx); free(dy);
    free(tmpx); free(tmpy); free(tmp);
    sb_free(&c);
}

/* ------------------------------------------------------------------------ */
/* FP64 evaluation of the same Joseph operator pair                         */
/* ------------------------------------------------------------------------ */

static inline double texel64(const double *img, int w, int h, int i, int j)
{
    return (i >= 0 && i < w && j >= 0 && j < h) ? img[(size_t)j * w + i] : 0.0;
}

/* ---- the line model (a second projection model; no reference counterpart) ----
 *
 * Ray s at angle theta is the line X cos - Y sin = t_s, t_s = s + 1/2 - axis,
 * with (cos, sin) the LUT pair scaled to unit length (the float LUT pair is a
 * unit vector only to ~3e-8) and a component below 2^-20 taken as 0: the
 * float LUT holds cos(pi/2) as -4.4e-8, the rounding of an exact zero, and a
 * ramp that narrow would make the weight of a ray along a pixel edge depend on
 * the last bits of d:
 * the geometry of the reference's BP (projector-parallel-joseph.cl:174-199,
 * u = X cos - Y sin + axis).  Its weight on pixel (x, y), centre (X, Y) =
 * (x + 1/2 - N/2, y + 1/2 - N/2), is the length of that line inside the unit
 * square: with d = t_s - (X cos - Y sin), M = max(|cos|, |sin|) and
 * m = min(|cos|, |sin|), a trapezoid in d of height 1/M, flat for
 * |d| <= (M - m)/2 and zero from |d| >= (M + m)/2.  For m = 0 (theta = 0) the
 * pixel spans d in [-1/2, 1/2) so that a ray along an edge counts once.
 * Every ray is active (no inactive-detector rule), and the BP is the exact
 * transpose of the FP. */
void orc64_line_dir(float cos_lut, float sin_lut, double *c, double *s)
{
    const double r = hypot((double)cos_lut, (double)sin_lut);
    double uc = cos_lut / r, us = sin_lut / r;
    const double tiny = 1.0 / 1048576.0;
    if (fabs(uc) < tiny) {
        uc = 0.0;
        us = us < 0.0 ? -1.0 : 1.0;
    } else if (fabs(us) < tiny) {
        us = 0.0;
        uc = uc < 0.0 ? -1.0 : 1.0;
    }
    *c = uc;
    *s = us;
}

double orc64_line_chord(double d, double c, double s)
{
    const double ac = fabs(c), as = fabs(s);
    const double M = ac > as ? ac : as, m = ac > as ? as : ac;
    const double h = 1.0 / M;
    if (m == 0.0)
        return (d >= -0.5 && d < 0.5) ? h : 0.0;
    double v = (0.5 * (M + m) - fabs(d)) / m;
    if (v <= 0.0)
        return 0.0;
    if (v > 1.0)
        v = 1.0;
    return h * v;
}

/* FP of the line model, ray by ray: along the ray's major axis, the (at most
 * three) pixels of each line whose square the ray can cross. */
static void fp64_line(const orc_geom *g, const double *vol, double *sino)
{
    const int N = g->n, D = g->n_dets, A = g->n_angles;
    const double axis = g->axis, h = 0.5 * N;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ang = 0; ang < A; ++ang) {
        for (int s = 0; s < D; ++s) {
            double c, sn;
            orc64_line_dir(g->cos_lut[ang], g->sin_lut[ang], &c, &sn);
            const double t = (double)s + 0.5 - axis;
            double acc = 0.0;
            if (fabs(sn) > fabs(c)) { /* columns: Y on the ray = (X cos - t) / sin */
                for (int x = 0; x < N; ++x) {
                    const double X = x + 0.5 - h, Yr = (X * c - t) / sn;
                    const int y0 = (int)floor(Yr + h - 0.5) - 1;
                    for (int y = y0; y < y0 + 4; ++y) {
                        if (y < 0 || y >= N)
                            continue;
                        const double Y = y + 0.5 - h;
                        acc += orc64_line_chord(t - (X * c - Y * sn), c, sn) * vol[(size_t)y * N + x];
                    }
                }
            } else { /* rows: X on the ray = (t + Y sin) / cos */
                for (int y = 0; y < N; ++y) {
                    const double Y = y + 0.5 - h, Xr = (t + Y * sn) / c;
                    const int x0 = (int)floor(Xr + h - 0.5) - 1;
                    for (int x = x0; x < x0 + 4; ++x) {
                        if (x < 0 || x >= N)
                            continue;
                        const double X = x + 0.5 - h;
                        acc += orc64_line_chord(t - (X * c - Y * sn), c, sn) * vol[(size_t)y * N + x];
                    }
                }
            }
            sino[(size_t)ang * D + s] = acc;
        }
    }
}

/* BP of the line model, pixel by pixel: the (at most three) bins of each
 * angle whose ray crosses the pixel; the same weights as fp64_line. */
static void bp64_line(const orc_geom *g, double *vol, const double *sino)
{
    const int N = g->n, D = g->n_dets, A = g->n_angles;
    const double axis = g->axis, h = 0.5 * N;
#pragma omp parallel for collapse(2) schedule(static)
    for (int y = 0; y < N; ++y) {
        for (int x = 0; x < N; ++x) {
            const double X = x + 0.5 - h, Y = y + 0.5 - h;
            double acc = 0.0;
            for (int ang = 0; ang < A; ++ang) {
                double c, sn;
                orc64_line_dir(g->cos_lut[ang], g->sin_lut[ang], &c, &sn);
                const double tc = X * c - Y * sn;
                const int s0 = (int)floor(tc + axis - 0.5) - 1;
                for (int s = s0; s < s0 + 4; ++s) {
                    if (s < 0 || s >= D)
                        continue;
                    acc += orc64_line_chord(((double)s + 0.5 - axis) - tc, c, sn) * sino[(size_t)ang * D + s];
                }
            }
            vol[(size_t)y * N + x] = acc;
        }
    }
}

/* Same rays as orc_fp_subset (including the half-pixel offsets of the
 * reference's start coordinate) but v_i = v0 + i * step evaluated in double.
 * With g->model = 1, the line model above. */
void orc64_fp(const orc_geom *g, const double *vol, double *sino)
{
    if (g->model == 1) {
        fp64_line(g, vol, sino);
        return;
    }
    const int N = g->n, D = g->n_dets, A = g->n_angles;
    const double axis = g->axis;
    const double required_width = 2.0 * axis;
    const double diff = (double)D - required_width;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ang = 0; ang < A; ++ang) {
        for (int s = 0; s < D; ++s) {
            if ((diff < 0 && (double)s < fabs(diff)) || (diff > 0 && (double)s > required_width)) {
                sino[(size_t)ang * D + s] = 0.0;
                continue;
            }
            const double st = g->sin_lut[ang], ct = g->cos_lut[ang];
            const int vertical = fabs(st) <= fabs(ct);
            const double t = (double)s + 0.5 - axis;
            double acc = 0.0;
            if (!vertical) {
                const double det_step = -1.0 / st, step = ct / st;
                const double v0 = t * det_step - 0.5 * N * step + 0.5 * N;
                for (int i = 0; i < N; ++i) {
                    double v = v0 + i * step - 0.5;
                    double fv = floor(v);
                    int j0 = (int)fv;
                    double b = v - fv;
                    acc += (1.0 - b) * texel64(vol, N, N, i, j0) + b * texel64(vol, N, N, i, j0 + 1);
                }
            } else {
                const double det_step = 1.0 / ct, step = st / ct;
                const double u0 = t * det_step - 0.5 * N * step + 0.5 * N;
                for (int j = 0; j < N; ++j) {
                    double u = u0 + j * step - 0.5;
                    double fu = floor(u);
                    int i0 = (int)fu;
                    double a = u - fu;
                    acc += (1.0 - a) * texel64(vol, N, N, i0, j) + a * texel64(vol, N, N, i0 + 1, j);
                }
            }
            sino[(size_t)ang * D + s] = acc;
        }
    }
}

/* Same as orc_bp over all angles: u = X cos - Y sin + axis, clamped to the
 * active detector window, linear interpolation, all in double. */
void orc64_bp(const orc_geom *g, double *vol, const double *sino)
{
    if (g->model == 1) {
        bp64_line(g, vol, sino);
        return;
    }
    const int N = g->n, D = g->n_dets, A = g->n_angles;
    const double axis = g->axis;
    const double diff = (double)D - 2.0 * axis;
    const double half_active = diff < 0 ? (double)D - axis : axis;
    const double edge0 = axis - half_active + 0.5, edge1 = axis + half_active - 0.5;
#pragma omp parallel for collapse(2) schedule(static)
    for (int py = 0; py < N; ++py) {
        for (int px = 0; px < N; ++px) {
            const double X = px + 0.5 - 0.5 * N, Y = py + 0.5 - 0.5 * N;
            double acc = 0.0;
            for (int ang = 0; ang < A; ++ang) {
                double u = X * (double)g->cos_lut[ang] - Y * (double)g->sin_lut[ang] + axis;
                if (u > edge1)
                    u = edge1;
                else if (u < edge0)
                    u = edge0;
                double w = u - 0.5, fw = floor(w);
                int i0 = (int)fw;
                double a = w - fw;
                const double *row = sino + (size_t)ang * D;
                double t0 = (i0 >= 0 && i0 < D) ? row[i0] : 0.0;
                double t1 = (i0 + 1 >= 0 && i0 + 1 < D) ? row[i0 + 1] : 0.0;
                acc += (1.0 - a) * t0 + a * t1;
            }
            vol[(size_t)py * N + px] = acc;
        }
    }
}

static void sirt64_core(const orc_geom *g, const double *b, double *x, uint32_t iters, double relax, int posc,
                        double r_min, int warm)
{
    const size_t nv = (size_t)g->n * g->n, ns = (size_t)g->n_angles * g->n_dets;
    double *R = (double *)malloc(ns * sizeof(double));
    double *C = (double *)malloc(nv * sizeof(double));
    double *ones_v = (double *)malloc(nv * sizeof(double));
    double *ones_s = (double *)malloc(ns * sizeof(double));
    double *r = (double *)malloc(ns * sizeof(double));
    double *bp = (double *)malloc(nv * sizeof(double));
    for (size_t i = 0; i < nv; ++i) ones_v[i] = 1.0;
    for (size_t i = 0; i < ns; ++i) ones_s[i] = 1.0;
    orc64_fp(g, ones_v, R);
    for (size_t i = 0; i < ns; ++i) R[i] = (R[i] != 0 && R[i] >= r_min) ? 1.0 / R[i] : 0.0;
    orc64_bp(g, C, ones_s);
    /* the line model takes no pixel covered by less than ORC_LINE_PIX_MIN of
     * chord length (a corner touch) */
    const double cmin = g->model == 1 ? ORC_LINE_PIX_MIN : 0.0;
    for (size_t i = 0; i < nv; ++i) C[i] = (C[i] != 0 && C[i] >= cmin) ? 1.0 / C[i] : 0.0;
    if (!warm)
        for (size_t i = 0; i < nv; ++i) x[i] = 0.0;
    for (uint32_t it = 0; it < iters; ++it) {
        orc64_fp(g, x, r);
        for (size_t i = 0; i < ns; ++i) r[i] = (b[i] - r[i]) * R[i];
        orc64_bp(g, bp, r);
        for (size_t i = 0; i < nv; ++i) {
            x[i] += relax * bp[i] * C[i];
            if (posc && x[i] < 0) x[i] = 0;
        }
    }
    free(R); free(C); free(ones_v); free(ones_s); free(r); free(bp);
}

void orc64_sirt(const orc_geom *g, const double *b, double *x, uint32_t iters, double relax, int posc,
                double r_min)
{
    sirt64_core(g, b, x, iters, relax, posc, r_min, 0);
}

/* single-angle FP/BP helpers for SART64 */
static void fp64_angle(const orc_geom *g, const double *vol, double *row, int ang)
{
    orc_geom g1 = *g;
    g1.n_angles = 1;
    g1.sin_lut = g->sin_lut + ang;
    g1.cos_lut = g->cos_lut + ang;
    orc64_fp(&g1, vol, row);
}

static void bp64_angle_add(const orc_geom *g, double *vol, double *tmp, const double *row, int ang, double relax)
{
    orc_geom g1 = *g;
    g1.n_angles = 1;
    g1.sin_lut = g->sin_lut + ang;
    g1.cos_lut = g->cos_lut + ang;
    const size_t nv = (size_t)g->n * g->n;
    orc64_bp(&g1, tmp, row);
    for (size_t i = 0; i < nv; ++i) vol[i] += relax * tmp[i];
}

static void sart64_core(const orc_geom *g, const double *b, double *x, uint32_t iters, double relax, double r_min,
                        int warm)
{
    const size_t nv = (size_t)g->n * g->n, ns = (size_t)g->n_angles * g->n_dets;
    const int D = g->n_dets;
    double *R = (double *)malloc(ns * sizeof(double));
    double *ones_v = (double *)malloc(nv * sizeof(double));
    double *row = (double *)malloc((size_t)D * sizeof(double));
    double *tmp = (double *)malloc(nv * sizeof(double));
    for (size_t i = 0; i < nv; ++i) ones_v[i] = 1.0;
    orc64_fp(g, ones_v, R);
    for (size_t i = 0; i < ns; ++i) R[i] = (R[i] != 0 && R[i] >= r_min) ? 1.0 / R[i] : 0.0;
    if (!warm)
        for (size_t i = 0; i < nv; ++i) x[i] = 0.0;
    for (uint32_t it = 0; it < iters; ++it) {
        for (int a = 0; a < g->n_angles; ++a) {
            fp64_angle(g, x, row, a);
            for (int s = 0; s < D; ++s)
                row[s] = (b[(size_t)a * D + s] - row[s]) * R[(size_t)a * D + s];
            bp64_angle_add(g, x, tmp, row, a, relax);
        }
    }
    free(R); free(ones_v); free(row); free(tmp);
}

void orc64_sart(const orc_geom *g, const double *b, double *x, uint32_t iters, double relax)
{
    sart64_core(g, b, x, iters, relax, 0.0, 0);
}

/* The smoothed TV gradient of ufo-math-tvstd-method.cl:5-57 (eps 1e-3, zero
 * outside the image) in FP64: the gradient of
 *   TV(f) = sum_{x,y} sqrt(eps + (f[x,y] - f[x-1,y])^2 + (f[x,y] - f[x,y+1])^2)
 * over the image extended by one zero pixel on each side. */
void orc64_l1_grad(const double *in, double *out, int n)
{
    const double eps = 0.001;
#define V(xx, yy) (((xx) < 0 || (yy) < 0 || (xx) >= n || (yy) >= n) ? 0.0 : in[(size_t)(yy) * n + (xx)])
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const double f3 = V(x, y), f1 = V(x, y - 1), f0 = V(x - 1, y - 1), f2 = V(x - 1, y);
            const double f4 = V(x + 1, y), f6 = V(x + 1, y + 1), f5 = V(x, y + 1);
            out[(size_t)y * n + x] = (2.0 * f3 - f2 - f5) / sqrt(eps + (f3 - f2) * (f3 - f2) + (f3 - f5) * (f3 - f5))
                                   - (f4 - f3) / sqrt(eps + (f4 - f3) * (f4 - f3) + (f4 - f6) * (f4 - f6))
                                   - (f1 - f3) / sqrt(eps + (f1 - f3) * (f1 - f3) + (f1 - f0) * (f1 - f0));
        }
#undef V
}

static double norm64(const double *v, size_t n, int l2)
{
    double s = 0;
    for (size_t i = 0; i < n; ++i)
        s += l2 ? v[i] * v[i] : fabs(v[i]);
    return l2 ? sqrt(s) : s;
}

/* ASD-POCS in FP64 with exact coordinates, both variants of orc_asdpocs.
 * Ray weights below r_min are zeroed as in orc64_sirt. */
void orc64_asdpocs(const orc_geom *g, const double *b, double *out, uint32_t iters, const orc_asdpocs_params *p,
                   double r_min, double *diag)
{
    const size_t nv = (size_t)g->n * g->n, ns = (size_t)g->n_angles * g->n_dets;
    const int tb = p->textbook;
    double *x = (double *)calloc(nv, sizeof(double));
    double *x_prev = (double *)calloc(nv, sizeof(double));
    double *x_res = (double *)malloc(nv * sizeof(double));
    double *b_res = (double *)malloc(ns * sizeof(double));
    double *grad = (double *)malloc(nv * sizeof(double));
    double dp = 1, dd = 1, dg = 1, dtgv = 1;
    double beta = p->beta, df_relax = p->df_relax_initial;
    const uint32_t ng = tb ? p->ng : 20;
    for (uint32_t it = 0; it < iters; ++it) {
        if (beta >= 0.0 && beta <= 1.0)
            df_relax = beta;
        if (tb)
            memcpy(x_prev, x, nv * sizeof(double));
        if (p->df_is_sirt)
            sirt64_core(g, b, x, p->df_iters, df_relax, 0, r_min, tb);
        else
            sart64_core(g, b, x, p->df_iters, df_relax, r_min, tb);
        if (p->positive_constraint)
            for (size_t i = 0; i < nv; ++i) x[i] = x[i] > 0 ? x[i] : 0.0;
        if (!tb)
            memcpy(out, x, nv * sizeof(double));
        orc64_fp(g, x, b_res);
        for (size_t i = 0; i < ns; ++i) b_res[i] = b[i] - b_res[i];
        dd = norm64(b_res, ns, tb);
        for (size_t i = 0; i < nv; ++i) x_res[i] = x[i] - x_prev[i];
        dp = norm64(x_res, nv, tb);
        if (it == 0)
            dtgv = p->alpha * dp;
        memcpy(x_prev, x, nv * sizeof(double));
        for (uint32_t k = 0; k < ng; ++k) {
            orc64_l1_grad(x, grad, g->n);
            const double norm = norm64(grad, nv, tb);
            if (tb && !(norm > 0))
                break;
            const double factor = dtgv / norm;
            for (size_t i = 0; i < nv; ++i) x[i] -= factor * grad[i];
        }
        for (size_t i = 0; i < nv; ++i) x_res[i] = x[i] - x_prev[i];
        dg = norm64(x_res, nv, tb);
        beta *= p->beta_red;
        if (diag) {
            diag[4 * it + 0] = dd;
            diag[4 * it + 1] = dp;
            diag[4 * it + 2] = dg;
            diag[4 * it + 3] = dtgv;
        }
        if (dg > p->r_max * dp && dd > 0.001)
            dtgv *= p->alpha_red;
        if (tb)
            memcpy(out, x, nv * sizeof(double));
    }
    free(x); free(x_prev); free(x_res); free(b_res); free(grad);
}

/* ---- split-Bregman TV with SIRT-preconditioned sweeps (FP64) ----
 * The outer loop of ufo-ir-sbtv-task.c:230-300 (u-subproblem, shrink,
 * Bregman update, update_db :360-402) with the u-subproblem
 *   min_u mu/2 |A u - f|^2_R + lambda/2 |e - D u|^2,  e = d - b,
 * approximated by `inner` sweeps u <- u + w P (mu A^T R (f - A u) + lambda D^T (e - D u)),
 * P = 1 / (mu A_n + 8 lambda), A_n = n_angles; differences as orc_dx .. orc_dyt. */
void orc64_sb_tv(const double *u, const double *ex, const double *ey, double *tv, int n)
{
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const int xl = x == 0 ? n - 1 : x - 1, xr = x == n - 1 ? 0 : x + 1;
            const int yu = y == 0 ? n - 1 : y - 1, yd = y == n - 1 ? 0 : y + 1;
            const size_t o = (size_t)y * n + x;
            /* (Dx^T h)[x] = h[x] - h[x+1], h = e - Dx u, (Dx u)[x] = u[x] - u[x-1] */
            const double hx0 = ex[o] - (u[o] - u[(size_t)y * n + xl]);
            const double hx1 = ex[(size_t)y * n + xr] - (u[(size_t)y * n + xr] - u[o]);
            const double hy0 = ey[o] - (u[o] - u[(size_t)yu * n + x]);
            const double hy1 = ey[(size_t)yd * n + x] - (u[(size_t)yd * n + x] - u[o]);
            tv[o] = (hx0 - hx1) + (hy0 - hy1);
        }
}

void orc64_sb_shrink(const double *u, double *bx, double *by, double *ex, double *ey, int n, double lambda)
{
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const int xl = x == 0 ? n - 1 : x - 1, yu = y == 0 ? n - 1 : y - 1;
            const size_t o = (size_t)y * n + x;
            const double gx = (u[o] - u[(size_t)y * n + xl]) + bx[o];
            const double gy = (u[o] - u[(size_t)yu * n + x]) + by[o];
            const double s = sqrt(gx * gx + gy * gy);
            const double t = fmax(s - 1.0 / lambda, 0.0) / fmax(s, 1e-12);
            const double dx = t * gx, dy = t * gy;
            bx[o] = gx - dx;
            by[o] = gy - dy;
            ex[o] = dx - bx[o];
            ey[o] = dy - by[o];
        }
}

void orc64_sbsirt(const orc_geom *g, const double *f, double *u, uint32_t outer, uint32_t inner, double relax,
                  double lambda, double mu, int posc, double r_min)
{
    const int n = g->n;
    const size_t nv = (size_t)n * n, ns = (size_t)g->n_angles * g->n_dets;
    double *R = (double *)calloc(ns, sizeof(double));
    double *ones_v = (double *)calloc(nv, sizeof(double));
    double *r = (double *)calloc(ns, sizeof(double));
    double *bp = (double *)calloc(nv, sizeof(double));
    double *tv = (double *)calloc(nv, sizeof(double));
    double *st = (double *)calloc(4 * nv, sizeof(double)); /* bx, by, ex, ey */
    double *P = (double *)calloc(nv, sizeof(double));
    if (!R || !ones_v || !r || !bp || !tv || !st || !P) {
        free(R); free(ones_v); free(r); free(bp); free(tv); free(st); free(P);
        return;
    }
    double *bx = st, *by = st + nv, *ex = st + 2 * nv, *ey = st + 3 * nv;
    for (size_t i = 0; i < nv; ++i) ones_v[i] = 1.0;
    orc64_fp(g, ones_v, R);
    for (size_t i = 0; i < ns; ++i) R[i] = (R[i] != 0 && R[i] >= r_min) ? 1.0 / R[i] : 0.0;
    if (g->model == 1) { /* per pixel: BP(1) of the model */
        for (size_t i = 0; i < ns; ++i) r[i] = 1.0;
        orc64_bp(g, bp, r);
        for (size_t i = 0; i < nv; ++i) {
            const double b1 = bp[i] >= ORC_LINE_PIX_MIN ? bp[i] : 0.0;
            P[i] = (b1 > 0 || lambda > 0) ? 1.0 / (mu * b1 + 8.0 * lambda) : 0.0;
        }
    } else {
        for (size_t i = 0; i < nv; ++i) P[i] = 1.0 / (mu * (double)g->n_angles + 8.0 * lambda);
    }
    for (size_t i = 0; i < nv; ++i) u[i] = 0.0;
    for (uint32_t k = 0; k < outer; ++k) {
        for (uint32_t j = 0; j < inner; ++j) {
            orc64_fp(g, u, r);
            for (size_t i = 0; i < ns; ++i) r[i] = (f[i] - r[i]) * R[i];
            orc64_bp(g, bp, r);
            orc64_sb_tv(u, ex, ey, tv, n);
            for (size_t i = 0; i < nv; ++i) {
                u[i] += relax * P[i] * (mu * bp[i] + lambda * tv[i]);
                if (posc && u[i] < 0) u[i] = 0;
            }
        }
        if (k + 1 < outer)
            orc64_sb_shrink(u, bx, by, ex, ey, n, lambda);
    }
    free(R); free(ones_v); free(r); free(bp); free(tv); free(st); free(P);
}
