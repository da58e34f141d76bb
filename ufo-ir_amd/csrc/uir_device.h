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

/* Reference-compatible Joseph FP (projector-parallel-joseph.cl:36-156) over
 * angles [ang0, ang0 + n_ang).  dir < 0 takes each angle's direction from the
 * LUT (|sin| <= |cos| -> vertical), else forces FP_vert (1) / FP_hor (0).
 * ray_w == NULL: w = r + corr * sum.  ray_w != NULL: w = (r + corr * sum) * ray_w
 * (the SIRT residual followed by operation_mul, ufo-ir-sirt-task.c:205-207). */
int uirk_fp_ref(const float *vol, const float *r_sino, float *w_sino, const float *ray_w,
                const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis,
                float corr, int ang0, int n_ang, int dir, void *stream);

/* Reference-compatible BP (projector-parallel-joseph.cl:158-218) over a list of
 * subsets (runs), one per-run accumulation each: w = r + relax * sum_run.
 * pix_w == NULL: accumulate mode (r_vol may equal w_vol).
 * pix_w != NULL: SIRT mode, tmp = sum_runs(relax * S_run) (from 0), then
 * w = tmp * pix_w + r (ufo-ir-sirt-task.c:208-212); posc clamps the result. */
int uirk_bp_ref(const float *r_vol, float *w_vol, const float *sino, int n_angles,
                const float *sin_lut, const float *cos_lut, int n, int n_dets, float axis,
                float relax, const uirk_runs *runs, const float *pix_w, int posc, void *stream);

/* Elementwise (ufo-ir-basic-ops.cl). */
int uirk_set(float *out, size_t n, float v, void *stream);
int uirk_inv(const float *in, float *out, size_t n, void *stream);
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

/* Finite differences (ufo-ir-gradient-processor.cl) and l1_grad. */
int uirk_diff(int kind, const float *in, float *out, uint32_t w, uint32_t h, void *stream);
enum { UIRK_DX = 0, UIRK_DXT = 1, UIRK_DY = 2, UIRK_DYT = 3 };
int uirk_l1_grad(const float *in, float *out, uint32_t w, uint32_t h, void *stream);

#ifdef __cplusplus
}
#endif
#endif
