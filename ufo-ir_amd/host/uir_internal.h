/*
 * uir_internal.h -- shared declarations of libuir's C host layer.
 */
#ifndef UIR_INTERNAL_H
#define UIR_INTERNAL_H

#include <hip/hip_runtime_api.h>
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>

#include "uir.h"
#include "uir_device.h"

#ifdef __cplusplus
extern "C" {
#endif

/* thread-local error message; returns s for tail calls */
uir_status uir_set_error(uir_status s, const char *fmt, ...);
uir_status uir_check_hip(int err, const char *what);

/* Reduction workspace owned by a resources object (or the per-device default). */
typedef struct {
    int device;
    void *d_ws;          /* device workspace */
    size_t ws_bytes;
    void *d_result;      /* 64 bytes device */
    void *h_result;      /* 64 bytes pinned host */
    pthread_mutex_t lock;
} uir_workspace;

struct uir_resources {
    int refcount;
    int device;
    hipStream_t stream;
    uir_workspace ws;
};

uir_status uir_workspace_reserve(uir_workspace *ws, size_t bytes);
uir_resources *uir_resources_retain(uir_resources *res);
void uir_workspace_release(uir_workspace *ws);

/* Reductions through a workspace on a stream; results returned on the host
 * (the stream is synchronised, as the reference's host-side reductions are). */
uir_status uir_ws_l1(uir_workspace *ws, const float *v, size_t n, void *stream, double *out);
uir_status uir_ws_dot_f64(uir_workspace *ws, const float *a, const float *b, size_t n, void *stream, double *out);
uir_status uir_ws_dot_compat(uir_workspace *ws, const float *a, const float *b, uint32_t w, uint32_t h,
                             void *stream, float *out);
uir_status uir_ws_normalize(uir_workspace *ws, float *v, size_t n, void *stream);

/* Device allocation with an optional failure-injection hook (tests). */
uir_status uir_dmalloc(void **ptr, size_t bytes);
void uir_dfree(void *ptr);

/* R = 1 / FP(1) of one geometry (FP(1) in FP64) into d_ray_w (A x D), 0 where
 * FP(1) = 0 or FP(1) < r_min; synchronises the stream (uir_batch.c). */
uir_status uir_ray_weights(void *stream, uint32_t n, uint32_t D, uint32_t A, float axis, const float *d_trig,
                           float r_min, float *d_ray_w);
/* The line model: a geometry's host tables (the unit directions replace the
 * LUT; runs of one direction family each), and its weights R = 1/FP(1) and
 * C = 1/BP(1) from the angle table on the device; synchronises the stream
 * (uir_batch.c). */
uir_status uir_line_geom(uint32_t A, uint32_t n, float axis, float *sin_lut, float *cos_lut, uirk_line_angle *h_line,
                         float *h_lk, uir_subset *runs, uint32_t *n_runs);
uir_status uir_line_weights(void *stream, uint32_t n, uint32_t D, uint32_t A, float r_min, const uir_subset *runs,
                            uint32_t n_runs, const uirk_line_angle *d_line, float *d_ray_w, float *d_pix_w);

/* A storage mode split into the sinogram and image element types; returns 0
 * for an unknown mode. */
static inline int uir_storage_split(int storage, int *sino, int *img)
{
    switch (storage) {
    case UIR_DTYPE_F32: *sino = *img = UIR_DTYPE_F32; return 1;
    case UIR_DTYPE_F16: *sino = *img = UIR_DTYPE_F16; return 1;
    case UIR_DTYPE_BF16: *sino = *img = UIR_DTYPE_BF16; return 1;
    case UIR_DTYPE_F16_SINO: *sino = UIR_DTYPE_F16; *img = UIR_DTYPE_F32; return 1;
    case UIR_DTYPE_BF16_SINO: *sino = UIR_DTYPE_BF16; *img = UIR_DTYPE_F32; return 1;
    default: return 0;
    }
}
static inline size_t uir_dtype_size(int dtype) { return dtype == UIR_DTYPE_F32 ? 4 : 2; }

/* autotuner (uir_tune.c): tune on a stream; the tile shape for a geometry */
uir_status uir_tune_run(void *stream, const uir_slice_spec *spec, uir_dtype storage, uint32_t reps,
                        uir_tune_result *out);
uirk_tile uir_tile_for(const uir_slice_spec *spec, uir_dtype storage);
/* Whether the geometry's fast FP uses single-value staging (wider windows):
 * when one angle step moves rays at the image edge by more than 15 pixels,
 * the pair window would cut its angle blocks to a few angles. */
int uir_fp_single_for(const uir_slice_spec *spec);

/* geometry helpers (ufo-ir-parallel-projector-task.c:436-498) */
void uir_geom_lut(uint32_t n_angles, float step, float *sin_lut, float *cos_lut);
uint32_t uir_geom_runs(const float *sin_lut, const float *cos_lut, uint32_t n_angles, uir_subset *out);
float uir_geom_axis(float axis_property, uint32_t n_dets);
int uir_geom_active_bins(uint32_t n_dets, float axis, int *lo, int *hi);
void uir_line_table(const float *sin_lut, const float *cos_lut, uint32_t n_angles, uint32_t n, float axis,
                    uirk_line_angle *out, float *trig_unit, float *fast_k);

#ifdef __cplusplus
}
#endif
#endif
