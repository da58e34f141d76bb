/*
 * uir.h -- C ABI of libuir, an MI355X-native (gfx950, HIP) implementation of
 * ufo-ir's iterative parallel-beam reconstruction path.
 *
 * The boundary replaces ufo-ir's task plugins (one GObject per task, driven by
 * ufo-core through UfoTaskIface) with plain C: opaque task handles created by
 * per-plugin factories, string-keyed properties with the reference names,
 * setup / get_requisition / process, and caller-owned device buffers.  No GLib,
 * no torch types.  Every entry point cites the reference interface it replaces
 * (paths relative to the ufo-ir source tree).
 *
 * Conventions
 *   - Every function returns uir_status; on failure uir_last_error() holds a
 *     thread-local message.
 *   - Buffers are device pointers (hipMalloc'd or uir_device_malloc'd) with
 *     UfoRequisition dims order: dims[0] = width (fastest), dims[1] = height.
 *     Images are [y][x]; sinograms are [angle][detector].
 *   - A task handle is thread-compatible, not thread-safe: one thread at a time
 *     per task, as one reference task instance is not reentrant
 *     (ufo-ir-parallel-projector-task.c:539-547 sets shared kernel args).
 */
#ifndef UIR_H
#define UIR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UIR_ABI_VERSION 1u

typedef enum {
    UIR_OK = 0,
    UIR_ERR_INVALID_ARGUMENT = 1, /* g_error("Incorrect volume size.") etc. in the reference */
    UIR_ERR_OUT_OF_MEMORY = 2,
    UIR_ERR_DEVICE = 3,           /* a HIP runtime error */
    UIR_ERR_UNSUPPORTED = 4,
    UIR_ERR_SETUP = 5,            /* GError UFO_TASK_ERROR_SETUP in the reference */
    UIR_ERR_NOT_FOUND = 6,        /* unknown plugin or property */
    UIR_ERR_BUSY = 7,
    UIR_ERR_CANCELLED = 8,
    UIR_ERR_TIMEOUT = 9,
    UIR_ERR_NO_DEVICE = 10,
    UIR_ERR_IO = 11               /* file could not be written */
} uir_status;

/* Element types.  The two *_SINO values are storage modes of the batched
 * engine, the server and the tuner: sinograms (measured data and residual)
 * in FP16 / BF16, images in FP32. */
typedef enum {
    UIR_DTYPE_F32 = 0,
    UIR_DTYPE_F16 = 1,
    UIR_DTYPE_BF16 = 2,
    UIR_DTYPE_F16_SINO = 3,
    UIR_DTYPE_BF16_SINO = 4
} uir_dtype;

uint32_t    uir_abi_version(void);
const char *uir_last_error(void);
const char *uir_status_string(uir_status s);

/* ------------------------------------------------------------------------ */
/* Device, memory, streams.  Replaces UfoResources / UfoGpuNode's command    */
/* queue (ufo-ir-sirt-task.c:170-171) and UfoBuffer's device allocations.    */
/* ------------------------------------------------------------------------ */
typedef struct uir_resources uir_resources;

uir_status uir_device_count(int *count);
/* One resources object per device: owns an in-order stream (the reference's
 * single in-order cl_command_queue per proc node) and device scratch. */
uir_status uir_resources_new(int device, uir_resources **out);
/* Drops the caller's reference; tasks set up on the resources keep their own,
 * so the stream lives until the last such task is released. */
void       uir_resources_free(uir_resources *res);
void      *uir_resources_stream(uir_resources *res);   /* hipStream_t */
int        uir_resources_device(const uir_resources *res);
uir_status uir_resources_synchronize(uir_resources *res);

uir_status uir_device_malloc(size_t bytes, void **ptr);
uir_status uir_device_free(void *ptr);
uir_status uir_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
uir_status uir_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
uir_status uir_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
uir_status uir_stream_synchronize(void *stream);
uir_status uir_mem_get_info(size_t *free_bytes, size_t *total_bytes);

/* Device event timing on a given stream (ms between two recorded points). */
typedef struct uir_timer uir_timer;
uir_status uir_timer_new(uir_timer **out);
void       uir_timer_free(uir_timer *t);
uir_status uir_timer_start(uir_timer *t, void *stream);
uir_status uir_timer_stop(uir_timer *t, void *stream);
uir_status uir_timer_elapsed_ms(uir_timer *t, float *ms);

/* ------------------------------------------------------------------------ */
/* Buffers and requisitions (UfoBuffer / UfoRequisition, 2-D float only).   */
/* ------------------------------------------------------------------------ */
typedef struct {
    void     *data;     /* device pointer, caller-owned */
    uint32_t  dims[2];  /* dims[0] = width, dims[1] = height */
    uir_dtype dtype;    /* tasks take UIR_DTYPE_F32 (the reference is FP32-only) */
} uir_buffer;

typedef struct {
    uint32_t n_dims;
    uint64_t dims[2];
} uir_requisition;

/* UfoIrProjectionsSubset, ufo-ir-parallel-projector-task.h:49-53 (12 bytes). */
typedef struct {
    uint32_t offset;
    uint32_t n;
    int32_t  direction; /* 1 = Vertical (|sin| <= |cos|), 0 = Horizontal */
} uir_subset;

/* UfoIrGeometryDims, ufo-ir-parallel-projector-task.h:55-61. */
typedef struct {
    uint64_t height;
    uint64_t width;
    uint64_t n_dets;
    uint64_t n_angles;
} uir_geometry_dims;

/* ------------------------------------------------------------------------ */
/* Tasks.  Factories replace the plugin constructors; properties keep the    */
/* reference names ('-' and '_' are interchangeable, as GLib canonicalises). */
/* ------------------------------------------------------------------------ */
typedef struct uir_task uir_task;

/* ufo_ir_parallel_projector_task_new, ufo-ir-parallel-projector-task.c:79-82.
 * Properties: axis_position (float, -1 = n_dets/2), step (float rad),
 * relaxation (float), correction_scale (float), is_forward (bool),
 * model (string, "joseph"), angles_num (uint, 0 = from sinogram)
 * (ufo-ir-projector-task.c:68-93, ufo-ir-state-dependent-task.c:75-80,
 *  ufo-ir-parallel-projector-task.c:123-137).
 * model selects the projection model at setup (:360-388; the reference ships
 * "joseph"): "joseph", or "line", exact ray-pixel intersection lengths with
 * every ray active and the BP the transpose of the FP (no reference
 * counterpart); any other name fails setup with UIR_ERR_SETUP.  The method
 * tasks, subsets and the kernel timing take the projector's model, the
 * batched SIRT and the server take it per slice (uir_slice_spec.model);
 * batched ASD-POCS takes batches of one model.
 * Extension property image_size (uint, 0 = n_dets): the N x N image edge. */
uir_task *uir_ir_parallel_projector_task_new(void);
/* ufo_ir_sirt_task_new, ufo-ir-sirt-task.c:142-146.  Properties: projector
 * (object), num-iterations (uint >= 1, default 10, ufo-ir-method-task.c:74-85),
 * relaxation_factor (float [0,1], default 0.25, ufo-ir-sirt-task.c:72-77).
 * Extension: positive_constraint (bool, default false) clamps x >= 0 after
 * each update. */
uir_task *uir_ir_sirt_task_new(void);
/* ufo_ir_sart_task_new, ufo-ir-sart-task.c:146-150 (relaxation_factor :76-81). */
uir_task *uir_ir_sart_task_new(void);
/* ufo_ir_asdpocs_task_new, ufo-ir-asdpocs-task.c:370-374.  Properties beta,
 * beta_red, ng, alpha, alpha_red, r_max, positive_constraint, df_minimizer
 * (ufo-ir-asdpocs-task.c:104-158). */
uir_task *uir_ir_asdpocs_task_new(void);
/* ufo_ir_sbtv_task_new, ufo-ir-sbtv-task.c:209-212.  Properties lambda, mu
 * (ufo-ir-sbtv-task.c:90-101). */
uir_task *uir_ir_sbtv_task_new(void);
/* Factory by plugin name ("parallel-projector", "sirt", "sart", "asdpocs",
 * "sbtv"), as ufo-core's plugin manager resolves package "ir" + name. */
uir_task *uir_task_new(const char *plugin);

uir_task   *uir_task_ref(uir_task *task);
void        uir_task_unref(uir_task *task);
const char *uir_task_plugin(const uir_task *task);

/* g_object_set / g_object_get on the task's GParamSpecs; values outside a
 * property's range are rejected with UIR_ERR_INVALID_ARGUMENT and leave the
 * property unchanged (GLib's behaviour for out-of-range values). */
uir_status uir_task_set_float(uir_task *task, const char *name, float value);
uir_status uir_task_set_uint(uir_task *task, const char *name, uint32_t value);
uir_status uir_task_set_bool(uir_task *task, const char *name, int value);
uir_status uir_task_set_string(uir_task *task, const char *name, const char *value);
uir_status uir_task_set_object(uir_task *task, const char *name, uir_task *value);
/* g_object_class_find_property: the value type of a property (the JSON
 * loader uses it to convert numbers, as GValue transforms do). */
typedef enum { UIR_PROP_FLOAT = 0, UIR_PROP_UINT = 1, UIR_PROP_BOOL = 2, UIR_PROP_STRING = 3,
               UIR_PROP_OBJECT = 4 } uir_prop_type;
uir_status uir_task_property_type(const uir_task *task, const char *name, uir_prop_type *type);
uir_status uir_task_get_float(const uir_task *task, const char *name, float *value);
uir_status uir_task_get_uint(const uir_task *task, const char *name, uint32_t *value);
uir_status uir_task_get_bool(const uir_task *task, const char *name, int *value);
uir_status uir_task_get_string(const uir_task *task, const char *name, char *buf, size_t buf_len);
uir_status uir_task_get_object(const uir_task *task, const char *name, uir_task **value);

/* UfoTaskIface.setup (e.g. ufo-ir-sirt-task.c:148-161). */
uir_status uir_task_setup(uir_task *task, uir_resources *res);
/* UfoTaskIface.get_requisition (ufo-ir-method-task.c:240-252 ->
 * ufo-ir-parallel-projector-task.c:305-347): output (n_dets, n_angles) when the
 * projector is forward, else the (N, N) image. */
uir_status uir_task_get_requisition(uir_task *task, const uir_buffer *inputs, uir_requisition *out);
/* UfoTaskIface.process (e.g. ufo-ir-sirt-task.c:163-222); output is written,
 * inputs[0] is the sinogram (methods) or the projector's input.  The SBTV task
 * min-max normalises inputs[0] in place, as the reference does
 * (ufo-ir-sbtv-task.c:239). */
uir_status uir_task_process(uir_task *task, const uir_buffer *inputs, uir_buffer *output);

/* Timing of a method's kernels, no counterpart in the reference (for
 * measurement): for the sart task, one sweep of the launches process() makes
 * (ufo-ir-sart-task.c:208-220), with every stride-th angle's FP and BP
 * launches bracketed by events on the task's stream; returns their average
 * durations in microseconds.  Synchronises; output holds a partial result. */
uir_status uir_task_time_kernels(uir_task *task, const uir_buffer *inputs, uir_buffer *output, uint32_t stride,
                                 float *fp_us, float *bp_us);

/* ufo_ir_state_dependent_task_forward / _backward,
 * ufo-ir-state-dependent-task.c:99-137: accumulate into output
 * (out = out + correction_scale * FP(in), out = out + relaxation * BP(in)). */
uir_status uir_projector_forward(uir_task *projector, const uir_buffer *input, uir_buffer *output);
uir_status uir_projector_backward(uir_task *projector, const uir_buffer *input, uir_buffer *output);
/* ufo_ir_parallel_projector_subset_fp / _bp,
 * ufo-ir-parallel-projector-task.h:93-94 (.c:167-195). */
uir_status uir_parallel_projector_subset_fp(uir_task *projector, const uir_buffer *volume,
                                            uir_buffer *sinogram, const uir_subset *subset);
uir_status uir_parallel_projector_subset_bp(uir_task *projector, uir_buffer *volume,
                                            const uir_buffer *sinogram, const uir_subset *subset);
/* ufo_ir_parallel_projector_get_host_sin_vals / _cos_vals and
 * generate_full_subsets_list, ufo-ir-parallel-projector-task.c:155-166,462-498;
 * valid after get_requisition. */
uir_status uir_parallel_projector_get_luts(uir_task *projector, const float **sin_vals,
                                           const float **cos_vals, uint32_t *n_angles);
uir_status uir_parallel_projector_get_subsets(uir_task *projector, const uir_subset **subsets,
                                              uint32_t *n_subsets);

/* ASD-POCS per-outer-iteration diagnostics (dd, dp, dg, dtgv) of the last
 * process call, ufo-ir-asdpocs-task.c:478-503; n_rows = num-iterations. */
uir_status uir_asdpocs_get_diagnostics(uir_task *task, const float **rows4, uint32_t *n_rows);

/* ------------------------------------------------------------------------ */
/* Batched SIRT.  No counterpart in the reference, which reconstructs one    */
/* slice per process() call (ufo-ir-method-task.c:226-232); this runs the    */
/* same SIRT update (ufo-ir-sirt-task.c:202-215) on many slices of uneven    */
/* geometry at once: images and sinograms stay resident in FP32, FP16 or     */
/* BF16 storage, interpolation and accumulation are FP32, one FP and one BP  */
/* launch per iteration cover every slice through a work table, and the      */
/* iterations are replayed from a captured hipGraph.                         */
/* ------------------------------------------------------------------------ */
typedef struct uir_batch uir_batch;

typedef struct {
    uint32_t n;              /* image edge (N x N) */
    uint32_t n_dets;         /* detector bins */
    uint32_t n_angles;       /* angles (sinogram rows) */
    float    step;           /* angular step, rad (projector "step") */
    float    axis_position;  /* projector "axis_position" (-1 = n_dets / 2) */
    float    relaxation;     /* SIRT "relaxation_factor" */
    int32_t  positive_constraint;
    /* Ray weights R = 1 / FP(1) are set to 0 for rays with FP(1) < ray_weight_min,
     * as for rays that miss the image.  0 = the reference's inv (only FP(1) = 0
     * gives 0, ufo-ir-basic-ops.cl:34).  Rays that clip an image corner by a tiny
     * fraction of a pixel otherwise get weights up to 1/FP(1) ~ 1e4-1e5 that
     * amplify noise into the corner pixels. */
    float    ray_weight_min;
    /* Projection model (the projector's "model"): 0 = joseph, 1 = line (exact
     * ray-pixel intersection lengths; its pixel weights 1 / BP(1) are an image
     * per geometry).  The batched SIRT and the server take both, side by
     * side; batched ASD-POCS takes either, one model per batch
     * (UIR_ERR_UNSUPPORTED for a mix). */
    int32_t  model;
} uir_slice_spec;

typedef struct {
    uint64_t device_bytes;       /* resident device memory of the batch */
    uint32_t n_slices, n_geometries;
    uint32_t fp_items, bp_items;
    uint32_t fp_grid, bp_grid;
    double   fp_samples, bp_samples; /* interpolation samples per iteration (in-image rays) */
} uir_batch_info;

uir_status uir_batch_new(uir_resources *res, uir_dtype storage, const uir_slice_spec *specs, uint32_t n_slices,
                         uir_batch **out);
void       uir_batch_free(uir_batch *batch);
/* Load slice `slice`'s sinogram ([n_angles][n_dets], FP32, host or device
 * memory) into the batch's storage dtype. */
uir_status uir_batch_set_sinogram(uir_batch *batch, uint32_t slice, const float *sino, int on_device);
/* x = 0 for every slice (ufo-ir-sirt-task.c:197). */
uir_status uir_batch_reset(uir_batch *batch);
/* `iterations` SIRT iterations on every slice; use_graph replays a captured
 * hipGraph, else launches directly.  Asynchronous on the batch stream. */
uir_status uir_batch_iterate(uir_batch *batch, uint32_t iterations, int use_graph);
/* Slice `slice`'s image as FP32 into host or device memory (synchronises). */
uir_status uir_batch_get_image(uir_batch *batch, uint32_t slice, float *dst, int on_device);
uir_status uir_batch_get_info(const uir_batch *batch, uir_batch_info *info);
/* Average device time (ms) of one iteration's FP launches and of its BP
 * launches (one launch per tile shape in use), over `reps` repetitions,
 * timed with events on the batch stream. */
uir_status uir_batch_time_kernels(uir_batch *batch, uint32_t reps, float *fp_ms, float *bp_ms);

/* Batched ASD-POCS: the asdpocs task (ufo-ir-asdpocs-task.c:411-509) with a
 * SART df-minimizer (ufo-ir-sart-task.c:173-228) on every slice of a batch at
 * once, fast-path projectors.  Per SART angle one FP and one BP launch cover
 * every slice that has the angle; the TV steps and the FP64 norm reductions
 * run on the device, and one iteration replays from a captured hipGraph.
 * compat = 1 is the reference's arithmetic (the df-minimizer restarts from
 * zero each iteration, L1 distances, 20 TV steps of length dtgv / |g|_1, the
 * output is the image before the TV steps); compat = 0 is ASD-POCS as
 * published (warm-started data step, L2 distances, ng steps along the
 * L2-normalised gradient, output after TV).  The slice spec's relaxation is
 * the df-minimizer's relaxation_factor until beta sets it; its
 * positive_constraint is not used (see positive_constraint below). */
typedef struct {
    float    beta, beta_red;      /* ufo-ir-asdpocs-task.c:104-116: 1.0, 0.995 */
    float    alpha, alpha_red;    /* :125-137: 0.2, 0.95 */
    float    r_max;               /* :139-144: 0.95 */
    uint32_t ng;                  /* TV steps per iteration, <= 1000 (compat: 20, as tvstd takes) */
    uint32_t df_iterations;       /* SART passes per iteration (the df-minimizer's num-iterations) */
    int32_t  positive_constraint; /* :146-151: 1 */
    int32_t  compat;              /* 1: the reference; 0: as published */
} uir_asdpocs_params;

void       uir_asdpocs_params_default(uir_asdpocs_params *p);
/* `iterations` ASD-POCS iterations of every slice from x = 0; the result
 * replaces the slice images (uir_batch_get_image).  Asynchronous on the batch
 * stream.  The first call allocates the method's own buffers (a second image
 * and an FP32 gradient per slice).  The batch's slices must share one
 * projection model (UIR_ERR_UNSUPPORTED otherwise). */
uir_status uir_batch_asdpocs(uir_batch *batch, uint32_t iterations, const uir_asdpocs_params *p, int use_graph);
/* Rows (dd, dp, dg, dtgv) of slice `slice` from the last uir_batch_asdpocs,
 * one per iteration (as uir_asdpocs_get_diagnostics); synchronises. */
uir_status uir_batch_asdpocs_diagnostics(uir_batch *batch, uint32_t slice, float *rows4, uint32_t max_rows,
                                         uint32_t *n_rows);

/* ------------------------------------------------------------------------ */
/* Tile-shape autotuner.  No counterpart in the reference (its work-group    */
/* shapes are the OpenCL runtime's).  The fast projector pair has variants: */
/* FP work items of 16, 8 or 32 angles (x 64 detectors), BP batches of 32  */
/* or 16 staged angles and tiles of 64 x 64 or 64 x 32 pixels.  Tuning times */
/* each on a probe of the geometry and keeps the fastest per geometry and    */
/* storage dtype; batches and the server then plan each geometry's work with */
/* its tuned shape (the default 16 / 32 / 64 when untuned).  Variants apply  */
/* the same arithmetic.  The BP's shape does not change any bit: a pixel's   */
/* window and coordinates are those of its 64 x 64 tile, and a batch or      */
/* server step whose BP table would not fill the device (4 work-groups per   */
/* CU) takes tiles of 32 rows whatever the tuned rows.  The FP's ray         */
/* coordinates are anchored per work item, so FP variants agree to FP32      */
/* rounding.                                                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t fp_angles;  /* 16, 8 or 32 */
    int32_t bp_angles;  /* 32 or 16 */
    int32_t bp_rows;    /* 64 or 32 */
} uir_tile;

typedef struct {
    uir_tile best;
    float    fp_ms[3];        /* fp_angles 16, 8, 32 */
    float    bp_ms[4];        /* (bp_angles, bp_rows) = (32, 64), (32, 32), (16, 64), (16, 32) */
    uint32_t probe_slices;    /* slices of the geometry in the probe */
} uir_tune_result;

/* Times every variant (`reps` launches each, 0 = 3) on the resources' stream
 * and records the fastest for this geometry and dtype. */
uir_status uir_tune_geometry(uir_resources *res, const uir_slice_spec *spec, uir_dtype storage, uint32_t reps,
                             uir_tune_result *out);
/* The recorded shape (UIR_ERR_NOT_FOUND, with the default in *out, if none). */
uir_status uir_tune_get(const uir_slice_spec *spec, uir_dtype storage, uir_tile *out);
/* Records a shape without timing (e.g. from an earlier run). */
uir_status uir_tune_set(const uir_slice_spec *spec, uir_dtype storage, const uir_tile *tile);
void       uir_tune_clear(void);

/* ------------------------------------------------------------------------ */
/* Reconstruction server: continuous batching.  No counterpart in the        */
/* reference (one slice per process() call in a ufo-core graph).  Requests   */
/* of any geometry arrive from any thread; a worker thread owning one device */
/* admits them in arrival order while the device budget and the slot limit   */
/* allow, and advances every admitted request by one iteration per step:     */
/* one FP and one BP launch over the work tables of all admitted requests.   */
/* A request leaves as soon as its iterations are done.  Each request's      */
/* arithmetic is independent of the others in the step, so its result is     */
/* bit-identical alone and in any mix.                                       */
/* ------------------------------------------------------------------------ */
typedef struct uir_server uir_server;

typedef enum {
    UIR_METHOD_SIRT = 0,      /* the SIRT update of ufo-ir-sirt-task.c:202-215 */
    UIR_METHOD_SBTV_SIRT = 1  /* split-Bregman TV (ufo-ir-sbtv-task.c's outer loop) with SIRT-preconditioned
                                 gradient sweeps for the u-subproblem in place of CGS */
} uir_method;

typedef struct {
    uir_slice_spec slice;      /* geometry, relaxation (SBTV_SIRT: in (0, 1]), positive_constraint */
    uint32_t method;           /* uir_method */
    uint32_t iterations;       /* SIRT: iterations; SBTV_SIRT: outer (Bregman) iterations */
    uint32_t inner_iterations; /* SBTV_SIRT: sweeps per outer iteration (>= 1) */
    float    lambda, mu;       /* SBTV_SIRT: ufo-ir-sbtv-task.c "lambda" and "mu" (> 0) */
} uir_request;

typedef struct {
    uir_dtype storage;         /* image / sinogram storage of every request */
    uint32_t  max_active;      /* most requests in one step (0 = no limit) */
    uint64_t  hbm_budget;      /* device bytes the server may hold (0 = free HBM at creation - hbm_reserve) */
    uint64_t  hbm_reserve;     /* bytes left free when hbm_budget = 0 */
    int32_t   autotune;        /* tune each new geometry's tile shape on first use (adds a probe) */
} uir_server_config;

typedef struct {
    double   t_submit, t_admit, t_done; /* seconds on uir_clock() */
    uint64_t device_bytes;              /* what the request held while admitted */
    uint32_t steps;                     /* steps it took part in (= its iterations) */
    uint32_t max_batch;                 /* most requests that shared one of its steps */
} uir_request_stats;

typedef struct {
    uint64_t budget, used;              /* device bytes */
    uint32_t active, pending;
    uint64_t steps, completed;
    double   busy_s;                    /* worker time with at least one request admitted */
    double   slice_iterations;          /* sum over steps of requests in the step */
} uir_server_info;

double     uir_clock(void);  /* monotonic seconds */
/* Starts the worker thread on `device`. */
uir_status uir_server_new(int device, const uir_server_config *cfg, uir_server **out);
/* Copies the request and its FP32 sinogram ([n_angles][n_dets], host memory)
 * and queues it.  A request whose footprint exceeds the whole budget is
 * refused with UIR_ERR_OUT_OF_MEMORY; bad parameters with
 * UIR_ERR_INVALID_ARGUMENT. */
uir_status uir_server_submit(uir_server *s, const uir_request *req, const float *sinogram, uint64_t *ticket);
/* Waits up to timeout_s (< 0: forever) for the request, copies its N x N FP32
 * image to host memory `image` (NULL: discard) and releases the ticket.
 * UIR_ERR_TIMEOUT leaves the ticket valid; a failed request returns its error. */
uir_status uir_server_wait(uir_server *s, uint64_t ticket, float *image, double timeout_s, uir_request_stats *stats);
uir_status uir_server_get_info(uir_server *s, uir_server_info *info);
/* Device bytes a request holds while admitted (its footprint for admission).
 * The request is validated as uir_server_submit validates it. */
uir_status uir_request_footprint(const uir_request *req, uir_dtype storage, uint64_t *bytes);
/* Cancels queued requests (their waits return UIR_ERR_CANCELLED), finishes
 * the admitted ones, joins the worker and frees everything.  No other thread
 * may be inside uir_server_submit / uir_server_wait on this server meanwhile. */
void       uir_server_free(uir_server *s);

/* ------------------------------------------------------------------------ */
/* Frame I/O and graph runner.  What ufo-core's "read" / "write" plugins and */
/* its JSON graph loader do for an ufo-ir graph (json-sample.json:1-56): one  */
/* 2-D FP32 frame per TIFF page, and a chain read -> ir tasks -> write.      */
/* ------------------------------------------------------------------------ */

/* Pages and the first page's size of a TIFF file. */
uir_status uir_tiff_info(const char *path, uint32_t *n_pages, uint32_t *width, uint32_t *height);
/* Page `page` as FP32 into host memory dst (capacity in floats; NULL dst:
 * size only).  Uncompressed 8/16/32-bit integer or 32/64-bit float samples. */
uir_status uir_tiff_read(const char *path, uint32_t page, float *dst, size_t capacity, uint32_t *width,
                         uint32_t *height);
/* Writes an FP32 page (append != 0: as a further page of an existing file). */
uir_status uir_tiff_write(const char *path, const float *data, uint32_t width, uint32_t height, int append);

typedef struct {
    uint32_t frames;        /* frames read, processed and written */
    uint32_t tasks;         /* ir tasks in the chain (nested ones not counted) */
    double   seconds;       /* wall time of the frame loop */
} uir_json_report;

/* Runs a ufo-core JSON graph whose edges form one chain: a "read" node
 * ("path": file, glob pattern or directory of TIFFs; optional "start",
 * "number"), ufo-ir task nodes (package "ir"; object-valued properties with a
 * "plugin" key become nested tasks, e.g. "projector", "df_minimizer"), and a
 * "write" node ("filename" with one printf integer conversion for one file
 * per frame, e.g. "out/out%03i.tif", or without one for a multi-page file).
 * Relative paths are taken from base_dir (NULL: the working directory).
 * Unknown properties are reported on stderr and skipped, as GLib does;
 * out-of-range values fail the run. */
uir_status uir_json_run(const char *json_text, const char *base_dir, int device, uir_json_report *report);

/* ------------------------------------------------------------------------ */
/* Kernel-level ABI (the reference's inner boundary).  All pointers are      */
/* device pointers; stream is a hipStream_t (NULL = default stream).         */
/* ------------------------------------------------------------------------ */

/* FP_hor / FP_vert, projector-parallel-joseph.cl:36-156 (args :37-45):
 * w_sino = r_sino + correction_scale * ray_sum for the subset's angles.
 * dims.width = dims.height = image edge, dims.n_dets = detectors.
 * r_sino may equal w_sino (the reference binds the same image). */
uir_status uir_fp_joseph(const float *volume, const float *r_sino, float *w_sino,
                         const float *sin_lut, const float *cos_lut, uir_geometry_dims dims,
                         float axis_pos, uir_subset part, float correction_scale, void *stream);
/* BP, projector-parallel-joseph.cl:158-218 (args :159-167):
 * w_vol = r_vol + relax * sum over the subset's angles. */
uir_status uir_bp_joseph(const float *r_volume, float *w_volume, const float *sinogram,
                         float relax, const float *sin_lut, const float *cos_lut,
                         uir_geometry_dims dims, float axis_pos, uir_subset part, void *stream);

/* Elementwise kernels, ufo-ir-basic-ops.cl:4-171,255-273 (wrappers
 * ufo-ir-basic-ops.h:28-84).  n = element count. */
uir_status uir_op_set(float *out, size_t n, float value, void *stream);                 /* operation_set */
uir_status uir_op_inv(const float *in, float *out, size_t n, void *stream);             /* operation_inv */
uir_status uir_op_mul(const float *a, const float *b, float *out, size_t n, void *stream); /* operation_mul */
uir_status uir_op_add(const float *a, const float *b, float *out, size_t n, void *stream); /* operation_add */
uir_status uir_op_deduction(const float *a, const float *b, float *out, size_t n, void *stream);
uir_status uir_op_deduction2(const float *a, const float *b, float m, float *out, size_t n, void *stream);
uir_status uir_op_add2(const float *a, const float *b, float m, float *out, size_t n, void *stream);
uir_status uir_op_mul_rows(const float *a, const float *b, float *out, uint32_t width,
                           uint32_t offset, uint32_t n_rows, void *stream);        /* op_mulRows */
uir_status uir_op_positive_constraint(const float *in, float *out, size_t n, void *stream); /* POSC */

/* Host-side ops of ufo-ir-basic-ops-processor.c moved to the device. */
uir_status uir_op_mul_scalar(float *v, size_t n, float m, void *stream);          /* :324-340 */
uir_status uir_op_sqrt(float *v, size_t n, void *stream);                         /* :415-430 */
uir_status uir_op_max_ew(const float *a, const float *b, float *out, size_t n, void *stream); /* :590-596 */
uir_status uir_op_div_ew(const float *a, const float *b, float *out, size_t n, void *stream); /* :598-604 */
uir_status uir_op_normalization(float *v, size_t n, void *stream);                /* :342-367 */

/* Reductions.  l1: ufo_ir_op_l1_norm (ufo-ir-basic-ops.c:259-277), summed in
 * FP64 on the device and rounded to float.  dot_compat / l2_compat reproduce
 * ufo_ir_basic_ops_processor_dot_product / _l2_norm bit for bit
 * (ufo-ir-basic-ops-processor.c:148-196,239-247), including its index pattern.
 * The result is written to host memory; the call synchronises the stream. */
uir_status uir_op_l1_norm(const float *v, uint32_t width, uint32_t height, float *result, void *stream);
uir_status uir_op_dot_compat(const float *a, const float *b, uint32_t width, uint32_t height,
                             float *result, void *stream);
uir_status uir_op_l2_compat(const float *a, uint32_t width, uint32_t height, float *result, void *stream);
uir_status uir_op_dot_f64(const float *a, const float *b, size_t n, double *result, void *stream);
/* uir_op_dot_compat without allocation or synchronisation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under hipGraph
 * capture, tight loops such as the split-Bregman CGS, harnesses that forbid
 * hipMalloc).  Same kernels and bit-identical result as uir_op_dot_compat,
 * written as a float to the caller's device memory d_result, enqueued on
 * stream and not waited for.  The row partials live in a 256 KiB device array
 * the library holds per device for the life of the process, so height is at
 * most 65536 (larger: UIR_ERR_INVALID_ARGUMENT); calls that overlap on one
 * device race on that array. */
uir_status uir_op_dot_compat_noalloc(const float *a, const float *b, uint32_t width, uint32_t height,
                                     float *d_result, void *stream);

/* Dx / Dxt / Dy / Dyt, ufo-ir-gradient-processor.cl:1-51 (periodic). */
uir_status uir_op_dx(const float *in, float *out, uint32_t width, uint32_t height, void *stream);
uir_status uir_op_dxt(const float *in, float *out, uint32_t width, uint32_t height, void *stream);
uir_status uir_op_dy(const float *in, float *out, uint32_t width, uint32_t height, void *stream);
uir_status uir_op_dyt(const float *in, float *out, uint32_t width, uint32_t height, void *stream);
/* l1_grad, ufo-math-tvstd-method.cl:5-57. */
uir_status uir_op_l1_grad(const float *in, float *out, uint32_t width, uint32_t height, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* UIR_H */
