/*
 * uir_core.c -- status/error handling, device memory, resources, timers and the
 * kernel-level C ABI (include/uir.h) over the HIP launchers.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "uir_internal.h"

static __thread char g_last_error[512];

uint32_t uir_abi_version(void) { return UIR_ABI_VERSION; }

const char *uir_last_error(void) { return g_last_error; }

uir_status uir_set_error(uir_status s, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
    va_end(ap);
    return s;
}

const char *uir_status_string(uir_status s)
{
    switch (s) {
    case UIR_OK: return "ok";
    case UIR_ERR_INVALID_ARGUMENT: return "invalid argument";
    case UIR_ERR_OUT_OF_MEMORY: return "out of memory";
    case UIR_ERR_DEVICE: return "device error";
    case UIR_ERR_UNSUPPORTED: return "unsupported";
    case UIR_ERR_SETUP: return "setup error";
    case UIR_ERR_NOT_FOUND: return "not found";
    case UIR_ERR_BUSY: return "busy";
    case UIR_ERR_CANCELLED: return "cancelled";
    case UIR_ERR_TIMEOUT: return "timeout";
    case UIR_ERR_NO_DEVICE: return "no device";
    case UIR_ERR_IO: return "i/o error";
    }
    return "unknown status";
}

uir_status uir_check_hip(int err, const char *what)
{
    if (err == hipSuccess)
        return UIR_OK;
    if (err == hipErrorOutOfMemory)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "%s: %s", what, hipGetErrorString((hipError_t)err));
    return uir_set_error(UIR_ERR_DEVICE, "%s: %s", what, hipGetErrorString((hipError_t)err));
}

/* ---- allocation with a failure-injection hook ---- */

static long g_fail_after = -1; /* UIR_FAULT_ALLOC_AFTER: fail the n-th allocation from now */

static void fault_init(void)
{
    static int done = 0;
    if (!done) {
        const char *e = getenv("UIR_FAULT_ALLOC_AFTER");
        if (e)
            g_fail_after = atol(e);
        done = 1;
    }
}

uir_status uir_dmalloc(void **ptr, size_t bytes)
{
    fault_init();
    *ptr = NULL;
    if (g_fail_after >= 0) {
        if (g_fail_after == 0)
            return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): injected allocation failure", bytes);
        --g_fail_after;
    }
    if (bytes == 0)
        bytes = 4;
    return uir_check_hip(hipMalloc(ptr, bytes), "hipMalloc");
}

void uir_dfree(void *ptr)
{
    if (ptr)
        (void)hipFree(ptr);
}

/* ---- devices, memory, streams ---- */

uir_status uir_device_count(int *count)
{
    if (!count)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "count is NULL");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return uir_set_error(UIR_ERR_NO_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    return UIR_OK;
}

uir_status uir_device_malloc(size_t bytes, void **ptr)
{
    if (!ptr)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "ptr is NULL");
    return uir_dmalloc(ptr, bytes);
}

uir_status uir_device_free(void *ptr)
{
    if (ptr)
        return uir_check_hip(hipFree(ptr), "hipFree");
    return UIR_OK;
}

uir_status uir_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream)
{
    uir_status s = uir_check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream), "hipMemcpyAsync H2D");
    if (s == UIR_OK)
        s = uir_check_hip(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    return s;
}

uir_status uir_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream)
{
    uir_status s = uir_check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream), "hipMemcpyAsync D2H");
    if (s == UIR_OK)
        s = uir_check_hip(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    return s;
}

uir_status uir_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream)
{
    return uir_check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream), "hipMemcpyAsync D2D");
}

uir_status uir_stream_synchronize(void *stream)
{
    return uir_check_hip(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

/* UIR_FAULT_FREE_HBM_BYTES overrides the free-memory figure (admission tests). */
uir_status uir_mem_get_info(size_t *free_bytes, size_t *total_bytes)
{
    size_t f = 0, t = 0;
    uir_status s = uir_check_hip(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    if (s != UIR_OK)
        return s;
    const char *e = getenv("UIR_FAULT_FREE_HBM_BYTES");
    if (e)
        f = (size_t)strtoull(e, NULL, 10);
    if (free_bytes)
        *free_bytes = f;
    if (total_bytes)
        *total_bytes = t;
    return UIR_OK;
}

/* ---- workspace ---- */

uir_status uir_workspace_reserve(uir_workspace *ws, size_t bytes)
{
    if (!ws->d_result) {
        uir_status s = uir_dmalloc(&ws->d_result, 64);
        if (s != UIR_OK)
            return s;
        s = uir_check_hip(hipHostMalloc(&ws->h_result, 64, hipHostMallocDefault), "hipHostMalloc");
        if (s != UIR_OK)
            return s;
    }
    if (bytes > ws->ws_bytes) {
        uir_dfree(ws->d_ws);
        ws->d_ws = NULL;
        ws->ws_bytes = 0;
        uir_status s = uir_dmalloc(&ws->d_ws, bytes);
        if (s != UIR_OK)
            return s;
        ws->ws_bytes = bytes;
    }
    return UIR_OK;
}

void uir_workspace_release(uir_workspace *ws)
{
    uir_dfree(ws->d_ws);
    uir_dfree(ws->d_result);
    if (ws->h_result)
        (void)hipHostFree(ws->h_result);
    ws->d_ws = ws->d_result = ws->h_result = NULL;
    ws->ws_bytes = 0;
}

static uir_status ws_fetch(uir_workspace *ws, size_t bytes, void *stream)
{
    uir_status s = uir_check_hip(hipMemcpyAsync(ws->h_result, ws->d_result, bytes, hipMemcpyDeviceToHost,
                                                (hipStream_t)stream), "hipMemcpyAsync result");
    if (s == UIR_OK)
        s = uir_check_hip(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    return s;
}

uir_status uir_ws_l1(uir_workspace *ws, const float *v, size_t n, void *stream, double *out)
{
    uir_status s = uir_workspace_reserve(ws, uirk_reduce_ws_bytes(n));
    if (s != UIR_OK)
        return s;
    s = uir_check_hip(uirk_l1_f64(v, n, (double *)ws->d_ws, (double *)ws->d_result, stream), "l1 reduction");
    if (s == UIR_OK)
        s = ws_fetch(ws, sizeof(double), stream);
    if (s == UIR_OK)
        *out = *(double *)ws->h_result;
    return s;
}

uir_status uir_ws_dot_f64(uir_workspace *ws, const float *a, const float *b, size_t n, void *stream, double *out)
{
    uir_status s = uir_workspace_reserve(ws, uirk_reduce_ws_bytes(n));
    if (s != UIR_OK)
        return s;
    s = uir_check_hip(uirk_dot_f64(a, b, n, (double *)ws->d_ws, (double *)ws->d_result, stream), "dot reduction");
    if (s == UIR_OK)
        s = ws_fetch(ws, sizeof(double), stream);
    if (s == UIR_OK)
        *out = *(double *)ws->h_result;
    return s;
}

uir_status uir_ws_dot_compat(uir_workspace *ws, const float *a, const float *b, uint32_t w, uint32_t h,
                             void *stream, float *out)
{
    uir_status s = uir_workspace_reserve(ws, uirk_reduce_ws_bytes((size_t)h));
    if (s != UIR_OK)
        return s;
    s = uir_check_hip(uirk_dot_compat(a, b, w, h, (float *)ws->d_ws, (float *)ws->d_result, stream), "dot_compat");
    if (s == UIR_OK)
        s = ws_fetch(ws, sizeof(float), stream);
    if (s == UIR_OK)
        *out = *(float *)ws->h_result;
    return s;
}

uir_status uir_ws_normalize(uir_workspace *ws, float *v, size_t n, void *stream)
{
    uir_status s = uir_workspace_reserve(ws, uirk_reduce_ws_bytes(n));
    if (s != UIR_OK)
        return s;
    s = uir_check_hip(uirk_minmax(v, n, (float *)ws->d_ws, (float *)ws->d_result, stream), "minmax");
    if (s == UIR_OK)
        s = uir_check_hip(uirk_normalize_apply(v, n, (const float *)ws->d_result, stream), "normalize");
    return s;
}

/* ---- resources ---- */

uir_status uir_resources_new(int device, uir_resources **out)
{
    if (!out)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = NULL;
    int n = 0;
    uir_status s = uir_device_count(&n);
    if (s != UIR_OK)
        return s;
    if (device < 0 || device >= n)
        return uir_set_error(UIR_ERR_NO_DEVICE, "device %d not present (%d devices)", device, n);
    s = uir_check_hip(hipSetDevice(device), "hipSetDevice");
    if (s != UIR_OK)
        return s;
    uir_resources *r = (uir_resources *)calloc(1, sizeof *r);
    if (!r)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
    r->refcount = 1;
    r->device = device;
    r->ws.device = device;
    pthread_mutex_init(&r->ws.lock, NULL);
    /* A blocking stream: it orders with the null stream, so null-stream copies
     * issued by callers see the results of work enqueued here. */
    s = uir_check_hip(hipStreamCreate(&r->stream), "hipStreamCreate");
    if (s != UIR_OK) {
        free(r);
        return s;
    }
    *out = r;
    return UIR_OK;
}

uir_resources *uir_resources_retain(uir_resources *r)
{
    if (r)
        __atomic_add_fetch(&r->refcount, 1, __ATOMIC_SEQ_CST);
    return r;
}

/* Drops the caller's reference; tasks set up on these resources hold their own. */
void uir_resources_free(uir_resources *r)
{
    if (!r)
        return;
    if (__atomic_sub_fetch(&r->refcount, 1, __ATOMIC_SEQ_CST) > 0)
        return;
    (void)hipStreamSynchronize(r->stream);
    uir_workspace_release(&r->ws);
    (void)hipStreamDestroy(r->stream);
    pthread_mutex_destroy(&r->ws.lock);
    free(r);
}

void *uir_resources_stream(uir_resources *r) { return r ? (void *)r->stream : NULL; }

int uir_resources_device(const uir_resources *r) { return r ? r->device : -1; }

uir_status uir_resources_synchronize(uir_resources *r)
{
    if (!r)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "resources is NULL");
    return uir_check_hip(hipStreamSynchronize(r->stream), "hipStreamSynchronize");
}

/* ---- timers ---- */

struct uir_timer {
    hipEvent_t start, stop;
};

uir_status uir_timer_new(uir_timer **out)
{
    uir_timer *t = (uir_timer *)calloc(1, sizeof *t);
    if (!t)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
    uir_status s = uir_check_hip(hipEventCreate(&t->start), "hipEventCreate");
    if (s == UIR_OK)
        s = uir_check_hip(hipEventCreate(&t->stop), "hipEventCreate");
    if (s != UIR_OK) {
        free(t);
        return s;
    }
    *out = t;
    return UIR_OK;
}

void uir_timer_free(uir_timer *t)
{
    if (!t)
        return;
    (void)hipEventDestroy(t->start);
    (void)hipEventDestroy(t->stop);
    free(t);
}

uir_status uir_timer_start(uir_timer *t, void *stream)
{
    return uir_check_hip(hipEventRecord(t->start, (hipStream_t)stream), "hipEventRecord");
}

uir_status uir_timer_stop(uir_timer *t, void *stream)
{
    return uir_check_hip(hipEventRecord(t->stop, (hipStream_t)stream), "hipEventRecord");
}

uir_status uir_timer_elapsed_ms(uir_timer *t, float *ms)
{
    uir_status s = uir_check_hip(hipEventSynchronize(t->stop), "hipEventSynchronize");
    if (s == UIR_OK)
        s = uir_check_hip(hipEventElapsedTime(ms, t->start, t->stop), "hipEventElapsedTime");
    return s;
}

/* ---- geometry helpers ---- */

/* create_lut_buffer, ufo-ir-parallel-projector-task.c:436-460 */
void uir_geom_lut(uint32_t n_angles, float step, float *sin_lut, float *cos_lut)
{
    for (uint32_t i = 0; i < n_angles; ++i) {
        double rad = (double)((float)i * step);
        sin_lut[i] = (float)sin(rad);
        cos_lut[i] = (float)cos(rad);
    }
}

/* generate_full_subsets_list, ufo-ir-parallel-projector-task.c:462-498 */
uint32_t uir_geom_runs(const float *s, const float *c, uint32_t n_angles, uir_subset *out)
{
    if (n_angles == 0)
        return 0;
    uint32_t k = 0;
    out[0].direction = fabsf(s[0]) <= fabsf(c[0]);
    out[0].offset = 0;
    out[0].n = 1;
    for (uint32_t i = 1; i < n_angles; ++i) {
        int d = fabsf(s[i]) <= fabsf(c[i]);
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

/* ufo-ir-parallel-projector-task.c:524-527,584-587 */
float uir_geom_axis(float axis_property, uint32_t n_dets)
{
    if (axis_property < 0)
        return (float)(n_dets / 2.0);
    return axis_property;
}

/* Constants of the line model per angle (uirk_line_angle) for an n x n image
 * and rotation axis `axis`: the unit direction from the float LUT pair,
 * components below 2^-20 taken as 0 (float cos(pi/2) is -4.4e-8, the
 * rounding of an exact zero); the FP position of detector s at sample i is
 * v = (t + Y s_) / c + n/2 - 1/2 (family 0, Y the centre of row i) or
 * (X c - t) / s_ + n/2 - 1/2 (family 1, X the centre of column i) with
 * t = s + 1/2 - axis, and the two taps at floor(v) and floor(v) + 1 lie at
 * ray offsets of magnitude M f and M (1 - f); the BP position of pixel (x, y)
 * is u = X c - Y s_ + axis - 1/2 with offsets f and 1 - f to bins floor(u)
 * and floor(u) + 1.
 * trig_unit (optional): (sin, cos) of the unit direction per angle, the
 * trig table of the fast path's line mode; fast_k (optional): its 8
 * constants per angle in terms of f - 1/2: the FP's phi = clamp(fp_k (f -
 * 1/2) + fp_b, 0, 1) (fp_k = M / m, fp_b = 1/2; axis-parallel: a step, fp_k
 * = 2^60, fp_b = 0 or 1 by the tie rule), 1/M, 0, and the BP's weights
 * clamp(bp_b0 - bp_k (f - 1/2), 0, 1/M), clamp(bp_b1 + bp_k (f - 1/2), 0,
 * 1/M) (bp_k = 1 / (M m), bp_b0 = bp_b1 = (M + m - 1) / (2 M m);
 * axis-parallel: bp_k = 2^60, bp_b0 = 1, bp_b1 = 0: f = 1/2 goes to the
 * lower bin), 1/M. */
void uir_line_table(const float *sin_lut, const float *cos_lut, uint32_t n_angles, uint32_t n, float axis,
                    uirk_line_angle *out, float *trig_unit, float *fast_k)
{
    const double h = 0.5 * n, tiny = 1.0 / 1048576.0;
    for (uint32_t a = 0; a < n_angles; ++a) {
        uirk_line_angle *L = &out[a];
        memset(L, 0, sizeof *L);
        const double r = hypot((double)cos_lut[a], (double)sin_lut[a]);
        double c = r > 0 ? cos_lut[a] / r : 1.0, s = r > 0 ? sin_lut[a] / r : 0.0;
        if (fabs(c) < tiny) {
            c = 0.0;
            s = s < 0 ? -1.0 : 1.0;
        } else if (fabs(s) < tiny) {
            s = 0.0;
            c = c < 0 ? -1.0 : 1.0;
        }
        const double M = fmax(fabs(c), fabs(s)), m = fmin(fabs(c), fabs(s));
        L->family = fabs(s) > fabs(c);
        L->aligned = m == 0.0;
        L->hm = (float)(1.0 / M);
        L->c = (float)c;
        L->s = (float)s;
        double step, tk;
        if (L->family == 0) {
            step = s / c;
            tk = 1.0 / c;
            L->tie0 = c < 0;
        } else {
            step = c / s;
            tk = -1.0 / s;
            L->tie0 = s > 0;
        }
        L->fp_step = (float)step;
        L->fp_tk = tk;
        L->fp_base = (0.5 - h) * step + h - 0.5 + (0.5 - (double)axis) * tk;
        L->bp_off = (float)((0.5 - h) * c + (double)axis - 0.5);
        if (!L->aligned) {
            const double q = m / M;
            L->fp_k = (float)(1.0 / q);
            L->fp_b0 = (float)((1.0 + q) / (2.0 * q));
            L->fp_b1 = (float)((q - 1.0) / (2.0 * q));
            L->bp_k = (float)(1.0 / (M * m));
            L->bp_b0 = (float)((M + m) / (2.0 * M * m));
            L->bp_b1 = (float)(((M + m) / 2.0 - 1.0) / (M * m));
        }
        if (trig_unit) {
            trig_unit[2 * a] = (float)s;
            trig_unit[2 * a + 1] = (float)c;
        }
        if (fast_k) {
            float *k = fast_k + 8 * (size_t)a;
            const float step = 1152921504606846976.0f; /* 2^60 */
            if (L->aligned) {
                k[0] = step;
                k[1] = L->tie0 ? 0.0f : 1.0f;
                k[4] = step;
                k[5] = 1.0f;
                k[6] = 0.0f;
            } else {
                k[0] = (float)(M / m);
                k[1] = 0.5f;
                k[4] = (float)(1.0 / (M * m));
                k[5] = k[6] = (float)((M + m - 1.0) / (2.0 * M * m));
            }
            k[2] = k[7] = L->hm;
            k[3] = 0.0f;
        }
    }
}

/* The active detector window of the back projection (projector-parallel-joseph.cl:
 * 174-182): u is clamped to [edge0, edge1], so the bins it interpolates lie in
 * [floor(edge0 - 0.5), floor(edge1 - 0.5)].  edge0 >= 0.5 and edge1 <= D - 0.5
 * for every axis, so the window lies inside [0, D - 1] when it is not empty;
 * it is empty for an axis below 0.5 or above D - 0.5 (the rotation axis
 * outside the first or last detector's centre).  Returns 1 for a non-empty
 * window, 0 otherwise. */
int uir_geom_active_bins(uint32_t n_dets, float axis, int *lo, int *hi)
{
    const float required_width = axis * 2;
    const float diff = (float)n_dets - required_width;
    const float half_active = diff < 0 ? (float)n_dets - axis : axis;
    const float edge0 = axis - half_active + 0.5f;
    const float edge1 = axis + half_active - 0.5f;
    const int l = (int)floorf(edge0 - 0.5f), h = (int)floorf(edge1 - 0.5f);
    if (lo)
        *lo = l;
    if (hi)
        *hi = h;
    return edge0 <= edge1 && l <= h && l >= 0 && h <= (int)n_dets - 1;
}

/* ---- kernel-level ABI ---- */

static uir_workspace g_default_ws[64];
static pthread_mutex_t g_default_ws_init = PTHREAD_MUTEX_INITIALIZER;
static int g_default_ws_ready[64];

static uir_workspace *default_ws(void)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        dev = 0;
    pthread_mutex_lock(&g_default_ws_init);
    if (!g_default_ws_ready[dev]) {
        memset(&g_default_ws[dev], 0, sizeof g_default_ws[dev]);
        g_default_ws[dev].device = dev;
        pthread_mutex_init(&g_default_ws[dev].lock, NULL);
        g_default_ws_ready[dev] = 1;
    }
    pthread_mutex_unlock(&g_default_ws_init);
    return &g_default_ws[dev];
}

static uir_status check_dims(uir_geometry_dims dims)
{
    if (dims.width == 0 || dims.height == 0 || dims.n_dets == 0)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "empty geometry");
    if (dims.width != dims.height)
        return uir_set_error(UIR_ERR_UNSUPPORTED, "image must be square (reference BP uses width for both axes, projector-parallel-joseph.cl:187-190)");
    if (dims.width > (1u << 20) || dims.n_dets > (1u << 20))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "geometry too large");
    return UIR_OK;
}

uir_status uir_fp_joseph(const float *volume, const float *r_sino, float *w_sino,
                         const float *sin_lut, const float *cos_lut, uir_geometry_dims dims,
                         float axis_pos, uir_subset part, float correction_scale, void *stream)
{
    uir_status s = check_dims(dims);
    if (s != UIR_OK)
        return s;
    if (!volume || !r_sino || !w_sino || !sin_lut || !cos_lut)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (part.direction != 0 && part.direction != 1)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "direction must be 0 or 1");
    /* a subset of a few angles is few rays: spread each ray over 16 threads
     * (bit-identical to one thread per ray) */
    if (part.n <= 4 && dims.width <= 16384)
        return uir_check_hip(uirk_fp_ref_split(volume, r_sino, w_sino, NULL, sin_lut, cos_lut, (int)dims.width,
                                               (int)dims.n_dets, axis_pos, correction_scale, (int)part.offset,
                                               (int)part.n, part.direction, stream), "FP launch");
    return uir_check_hip(uirk_fp_ref(volume, r_sino, w_sino, NULL, sin_lut, cos_lut, (int)dims.width,
                                     (int)dims.n_dets, axis_pos, correction_scale, (int)part.offset,
                                     (int)part.n, part.direction, stream), "FP launch");
}

uir_status uir_bp_joseph(const float *r_volume, float *w_volume, const float *sinogram,
                         float relax, const float *sin_lut, const float *cos_lut,
                         uir_geometry_dims dims, float axis_pos, uir_subset part, void *stream)
{
    uir_status s = check_dims(dims);
    if (s != UIR_OK)
        return s;
    if (!r_volume || !w_volume || !sinogram || !sin_lut || !cos_lut)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (dims.n_angles < (uint64_t)part.offset + part.n)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "subset exceeds n_angles");
    uirk_runs runs;
    memset(&runs, 0, sizeof runs);
    runs.count = 1;
    runs.offset[0] = part.offset;
    runs.n[0] = part.n;
    return uir_check_hip(uirk_bp_ref(r_volume, w_volume, sinogram, (int)dims.n_angles, sin_lut, cos_lut,
                                     (int)dims.width, (int)dims.n_dets, axis_pos, relax, &runs, NULL, 0,
                                     stream), "BP launch");
}

#define CHECK_PTR(p) do { if (!(p)) return uir_set_error(UIR_ERR_INVALID_ARGUMENT, #p " is NULL"); } while (0)

uir_status uir_op_set(float *out, size_t n, float value, void *stream)
{
    CHECK_PTR(out);
    return uir_check_hip(uirk_set(out, n, value, stream), "operation_set");
}

uir_status uir_op_inv(const float *in, float *out, size_t n, void *stream)
{
    CHECK_PTR(in); CHECK_PTR(out);
    return uir_check_hip(uirk_inv(in, out, n, stream), "operation_inv");
}

static uir_status binop(int op, const float *a, const float *b, float m, float *out, size_t n, void *stream, const char *name)
{
    CHECK_PTR(a); CHECK_PTR(b); CHECK_PTR(out);
    return uir_check_hip(uirk_binary(op, a, b, m, out, n, stream), name);
}

uir_status uir_op_mul(const float *a, const float *b, float *out, size_t n, void *stream)
{ return binop(UIRK_MUL, a, b, 0, out, n, stream, "operation_mul"); }
uir_status uir_op_add(const float *a, const float *b, float *out, size_t n, void *stream)
{ return binop(UIRK_ADD, a, b, 0, out, n, stream, "operation_add"); }
uir_status uir_op_deduction(const float *a, const float *b, float *out, size_t n, void *stream)
{ return binop(UIRK_SUB, a, b, 0, out, n, stream, "operation_deduction"); }
uir_status uir_op_deduction2(const float *a, const float *b, float m, float *out, size_t n, void *stream)
{ return binop(UIRK_SUB_SCALED, a, b, m, out, n, stream, "operation_deduction2"); }
uir_status uir_op_add2(const float *a, const float *b, float m, float *out, size_t n, void *stream)
{ return binop(UIRK_ADD_SCALED, a, b, m, out, n, stream, "operation_add2"); }
uir_status uir_op_max_ew(const float *a, const float *b, float *out, size_t n, void *stream)
{ return binop(UIRK_MAX, a, b, 0, out, n, stream, "max_element_wise"); }
uir_status uir_op_div_ew(const float *a, const float *b, float *out, size_t n, void *stream)
{ return binop(UIRK_DIV, a, b, 0, out, n, stream, "div_element_wise"); }

uir_status uir_op_mul_rows(const float *a, const float *b, float *out, uint32_t width,
                           uint32_t offset, uint32_t n_rows, void *stream)
{
    CHECK_PTR(a); CHECK_PTR(b); CHECK_PTR(out);
    return uir_check_hip(uirk_mul_rows(a, b, out, width, offset, n_rows, stream), "op_mulRows");
}

uir_status uir_op_positive_constraint(const float *in, float *out, size_t n, void *stream)
{
    CHECK_PTR(in); CHECK_PTR(out);
    return uir_check_hip(uirk_unary(UIRK_POSC, in, out, n, 0, stream), "POSC");
}

uir_status uir_op_mul_scalar(float *v, size_t n, float m, void *stream)
{
    CHECK_PTR(v);
    return uir_check_hip(uirk_unary(UIRK_SCALE, v, v, n, m, stream), "mul_scalar");
}

uir_status uir_op_sqrt(float *v, size_t n, void *stream)
{
    CHECK_PTR(v);
    return uir_check_hip(uirk_unary(UIRK_SQRT, v, v, n, 0, stream), "sqrt");
}

uir_status uir_op_normalization(float *v, size_t n, void *stream)
{
    CHECK_PTR(v);
    uir_workspace *ws = default_ws();
    pthread_mutex_lock(&ws->lock);
    uir_status s = uir_ws_normalize(ws, v, n, stream);
    if (s == UIR_OK)
        s = uir_stream_synchronize(stream);
    pthread_mutex_unlock(&ws->lock);
    return s;
}

uir_status uir_op_l1_norm(const float *v, uint32_t width, uint32_t height, float *result, void *stream)
{
    CHECK_PTR(v); CHECK_PTR(result);
    uir_workspace *ws = default_ws();
    double d = 0;
    pthread_mutex_lock(&ws->lock);
    uir_status s = uir_ws_l1(ws, v, (size_t)width * height, stream, &d);
    pthread_mutex_unlock(&ws->lock);
    if (s == UIR_OK)
        *result = (float)d;
    return s;
}

uir_status uir_op_dot_compat(const float *a, const float *b, uint32_t width, uint32_t height,
                             float *result, void *stream)
{
    CHECK_PTR(a); CHECK_PTR(b); CHECK_PTR(result);
    if (height == 0)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "height is 0");
    uir_workspace *ws = default_ws();
    pthread_mutex_lock(&ws->lock);
    uir_status s = uir_ws_dot_compat(ws, a, b, width, height, stream, result);
    pthread_mutex_unlock(&ws->lock);
    return s;
}

uir_status uir_op_l2_compat(const float *a, uint32_t width, uint32_t height, float *result, void *stream)
{
    float d = 0;
    uir_status s = uir_op_dot_compat(a, a, width, height, &d, stream);
    if (s == UIR_OK)
        *result = (float)sqrt(d);
    return s;
}

uir_status uir_op_dot_f64(const float *a, const float *b, size_t n, double *result, void *stream)
{
    CHECK_PTR(a); CHECK_PTR(b); CHECK_PTR(result);
    uir_workspace *ws = default_ws();
    pthread_mutex_lock(&ws->lock);
    uir_status s = uir_ws_dot_f64(ws, a, b, n, stream, result);
    pthread_mutex_unlock(&ws->lock);
    return s;
}

static uir_status diffop(int kind, const float *in, float *out, uint32_t w, uint32_t h, void *stream, const char *name)
{
    CHECK_PTR(in); CHECK_PTR(out);
    if (in == out)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: in and out must differ", name);
    return uir_check_hip(uirk_diff(kind, in, out, w, h, stream), name);
}

uir_status uir_op_dx(const float *in, float *out, uint32_t w, uint32_t h, void *stream)
{ return diffop(UIRK_DX, in, out, w, h, stream, "Dx"); }
uir_status uir_op_dxt(const float *in, float *out, uint32_t w, uint32_t h, void *stream)
{ return diffop(UIRK_DXT, in, out, w, h, stream, "Dxt"); }
uir_status uir_op_dy(const float *in, float *out, uint32_t w, uint32_t h, void *stream)
{ return diffop(UIRK_DY, in, out, w, h, stream, "Dy"); }
uir_status uir_op_dyt(const float *in, float *out, uint32_t w, uint32_t h, void *stream)
{ return diffop(UIRK_DYT, in, out, w, h, stream, "Dyt"); }

uir_status uir_op_l1_grad(const float *in, float *out, uint32_t w, uint32_t h, void *stream)
{
    CHECK_PTR(in); CHECK_PTR(out);
    if (in == out)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "l1_grad: in and out must differ");
    return uir_check_hip(uirk_l1_grad(in, out, w, h, stream), "l1_grad");
}
