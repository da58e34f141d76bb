/*
 * uir_tune.c -- per-geometry autotuner for the fast projector tile shapes.
 *
 * For one geometry and storage dtype it builds a probe of a few slices on
 * scratch device memory (the work a batch of that geometry does per
 * iteration), times every FP variant (16, 8 or 32 angles per work item) and every
 * BP variant (32 or 16 staged angles x tiles of 64 or 32 rows) with events on
 * the given stream, and keeps the fastest of each in a process-wide cache
 * keyed by geometry and dtype.  Batches and the server plan their work tables
 * with the cached shape of each geometry (the default shape if untuned).
 * Every variant applies the same arithmetic to each sample and pixel; the work
 * split moves the points where FP32 ray and pixel coordinates are anchored,
 * so results agree to FP32 rounding.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "uir_internal.h"

typedef struct {
    uint32_t n, D, A;
    float step, axis;
    int dtype;
    uir_tile tile;
} tune_entry;

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static tune_entry *g_cache;
static size_t g_n, g_cap;

static const uir_tile k_fp_variants[3] = {{16, 32, 64}, {8, 32, 64}, {32, 32, 64}};
static const uir_tile k_bp_variants[4] = {{16, 32, 64}, {16, 32, 32}, {16, 16, 64}, {16, 16, 32}};

static int key_eq(const tune_entry *e, const uir_slice_spec *s, int dtype)
{
    return e->n == s->n && e->D == s->n_dets && e->A == s->n_angles && e->step == s->step &&
           e->axis == uir_geom_axis(s->axis_position, s->n_dets) && e->dtype == dtype;
}

uir_status uir_tune_get(const uir_slice_spec *spec, uir_dtype storage, uir_tile *out)
{
    if (!spec || !out)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    const uir_tile def = {UIRK_FP_ANGLES, 32, UIRK_BP_TILE};
    *out = def;
    pthread_mutex_lock(&g_mu);
    for (size_t i = 0; i < g_n; ++i)
        if (key_eq(&g_cache[i], spec, storage)) {
            *out = g_cache[i].tile;
            pthread_mutex_unlock(&g_mu);
            return UIR_OK;
        }
    pthread_mutex_unlock(&g_mu);
    return uir_set_error(UIR_ERR_NOT_FOUND, "geometry not tuned (default tiles)");
}

uir_status uir_tune_set(const uir_slice_spec *spec, uir_dtype storage, const uir_tile *tile)
{
    if (!spec || !tile)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    const uirk_tile t = {tile->fp_angles, tile->bp_angles, tile->bp_rows, 0};
    if (!((t.fp_angles == 16 || t.fp_angles == 8 || t.fp_angles == 32) && (t.bp_angles == 32 || t.bp_angles == 16) &&
          (t.bp_rows == 64 || t.bp_rows == 32)))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "tile %d/%d/%d is not a variant", t.fp_angles, t.bp_angles,
                             t.bp_rows);
    pthread_mutex_lock(&g_mu);
    for (size_t i = 0; i < g_n; ++i)
        if (key_eq(&g_cache[i], spec, storage)) {
            g_cache[i].tile = *tile;
            pthread_mutex_unlock(&g_mu);
            return UIR_OK;
        }
    if (g_n == g_cap) {
        size_t c = g_cap ? 2 * g_cap : 16;
        tune_entry *ne = (tune_entry *)realloc(g_cache, c * sizeof *ne);
        if (!ne) {
            pthread_mutex_unlock(&g_mu);
            return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "realloc");
        }
        g_cache = ne;
        g_cap = c;
    }
    tune_entry *e = &g_cache[g_n++];
    e->n = spec->n;
    e->D = spec->n_dets;
    e->A = spec->n_angles;
    e->step = spec->step;
    e->axis = uir_geom_axis(spec->axis_position, spec->n_dets);
    e->dtype = storage;
    e->tile = *tile;
    pthread_mutex_unlock(&g_mu);
    return UIR_OK;
}

void uir_tune_clear(void)
{
    pthread_mutex_lock(&g_mu);
    free(g_cache);
    g_cache = NULL;
    g_n = g_cap = 0;
    pthread_mutex_unlock(&g_mu);
}

int uir_fp_single_for(const uir_slice_spec *spec)
{
    /* Measured at 1024^2 / 720 angles (step x N = 4.5) and 2048^2 / 900
     * (7.1) the pair FP is 4-15 % faster; at 1024^2 / 120 (27) and
     * 2048^2 / 90 (71) the single-value FP is 15-33 % faster. */
    return fabs((double)spec->step) * (double)spec->n > 15.0;
}

uirk_tile uir_tile_for(const uir_slice_spec *spec, uir_dtype storage)
{
    uir_tile t;
    (void)uir_tune_get(spec, storage, &t);
    const uirk_tile k = {t.fp_angles, t.bp_angles, t.bp_rows, uir_fp_single_for(spec)};
    return k;
}

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

/* Average ms of `reps` launches of one variant (after one untimed launch). */
static uir_status time_variant(hipStream_t st, int fp, int mode, const uirk_slice *d_sl, const void *d_items,
                               int n_items, const uirk_xcd_part *part, int sdt, int idt, const uirk_tile *tile,
                               int grid, int *d_err, uint32_t reps, float *ms)
{
    hipEvent_t e0, e1;
    uir_status s = uir_check_hip(hipEventCreate(&e0), "hipEventCreate");
    if (s != UIR_OK)
        return s;
    if ((s = uir_check_hip(hipEventCreate(&e1), "hipEventCreate")) != UIR_OK) {
        (void)hipEventDestroy(e0);
        return s;
    }
    for (uint32_t r = 0; r <= reps && s == UIR_OK; ++r) {
        if (r == 1)
            s = uir_check_hip(hipEventRecord(e0, st), "hipEventRecord");
        if (s != UIR_OK)
            break;
        const int e = fp ? uirk_fp_table(d_sl, NULL, (const uirk_fp_item *)d_items, n_items, part, mode, sdt, idt,
                                         tile, grid, d_err, st)
                         : uirk_bp_table(d_sl, NULL, (const uirk_bp_item *)d_items, n_items, part, mode, sdt, idt,
                                         tile, grid, d_err, st);
        s = uir_check_hip(e, fp ? "FP launch" : "BP launch");
    }
    if (s == UIR_OK)
        s = uir_check_hip(hipEventRecord(e1, st), "hipEventRecord");
    if (s == UIR_OK)
        s = uir_check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
    float t = 0;
    if (s == UIR_OK)
        s = uir_check_hip(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (s == UIR_OK)
        *ms = t / (float)reps;
    return s;
}

uir_status uir_tune_run(void *stream, const uir_slice_spec *spec, uir_dtype storage, uint32_t reps,
                        uir_tune_result *out)
{
    if (!spec || !out)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    int sdt = 0, idt = 0;
    if (!uir_storage_split(storage, &sdt, &idt))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "unknown storage dtype");
    const uint32_t n = spec->n, D = spec->n_dets, A = spec->n_angles;
    if (!n || !D || !A || n > 16384 || (uint64_t)n * n * 4 >= (1ull << 32) || (uint64_t)A * D * 4 >= (1ull << 32))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "bad geometry %ux%u, %u angles", n, D, A);
    if (reps == 0)
        reps = 3;
    memset(out, 0, sizeof *out);
    hipStream_t st = (hipStream_t)stream;
    const size_t ess = uir_dtype_size(sdt), esi = uir_dtype_size(idt);
    const size_t nv = (size_t)n * n, ns = (size_t)A * D;
    /* about 8 M pixels of probe, at least one slice, at most 16 */
    uint32_t B = (uint32_t)((8u << 20) / nv);
    B = B < 1 ? 1 : (B > 16 ? 16 : B);
    out->probe_slices = B;

    const float axis = uir_geom_axis(spec->axis_position, D);
    float *h_sin = (float *)malloc(sizeof(float) * A), *h_cos = (float *)malloc(sizeof(float) * A);
    float *trig = (float *)malloc(sizeof(float) * 2 * A);
    uir_subset *runs = (uir_subset *)malloc(sizeof(uir_subset) * A);
    uirk_slice *h_sl = (uirk_slice *)calloc(B, sizeof(uirk_slice));
    double *cost = NULL;
    void *mem = NULL, *d_items = NULL;
    uirk_slice *d_sl = NULL;
    int *d_err = NULL;
    uir_status s = UIR_OK;
    if (!h_sin || !h_cos || !trig || !runs || !h_sl) {
        s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
        goto out;
    }
    uir_geom_lut(A, spec->step, h_sin, h_cos);
    const uint32_t n_runs = uir_geom_runs(h_sin, h_cos, A, runs);
    uirk_plan_trig(h_sin, h_cos, A, trig);
    /* per slice: x (storage), b and r (storage); shared: ray weights (FP32, ones), trig */
    const size_t per = align_up(nv * esi) + 2 * align_up(ns * ess);
    const size_t total = (size_t)B * per + align_up(ns * 4) + align_up((size_t)A * 8);
    if ((s = uir_dmalloc(&mem, total)) != UIR_OK || (s = uir_dmalloc((void **)&d_err, sizeof(int))) != UIR_OK)
        goto out;
    /* content does not change the work: x = 0, b = 0, ray weights = 1 */
    if ((s = uir_check_hip(uirk_set((float *)mem, total / 4, 0.0f, st), "set")) != UIR_OK)
        goto out;
    char *base = (char *)mem;
    float *d_rw = (float *)(base + (size_t)B * per);
    float *d_trig = (float *)(base + (size_t)B * per + align_up(ns * 4));
    if ((s = uir_check_hip(uirk_set(d_rw, ns, 1.0f, st), "set")) != UIR_OK ||
        (s = uir_check_hip(hipMemcpyAsync(d_trig, trig, sizeof(float) * 2 * A, hipMemcpyHostToDevice, st),
                           "hipMemcpyAsync")) != UIR_OK ||
        (s = uir_check_hip(hipMemsetAsync(d_err, 0, sizeof(int), st), "hipMemsetAsync")) != UIR_OK)
        goto out;
    for (uint32_t i = 0; i < B; ++i) {
        uirk_slice *f = &h_sl[i];
        char *p = base + (size_t)i * per;
        f->x = p;
        f->x_out = p;
        f->b = p + align_up(nv * esi);
        f->r = p + align_up(nv * esi) + align_up(ns * ess);
        f->ray_w = d_rw;
        f->trig = d_trig;
        f->n = (int32_t)n;
        f->n_dets = (int32_t)D;
        f->n_angles = (int32_t)A;
        f->axis = axis;
        f->corr = -1.0f;
        f->relax = 0.25f;
        f->pix_scale = 1.0f / (float)A;
        f->n_runs = 1;
        f->run_off[0] = 0;
        f->run_n[0] = A;
    }
    /* FP and BP both read the descriptors: FP's b / r and BP's b (= r) agree here */
    if ((s = uir_dmalloc((void **)&d_sl, B * sizeof(uirk_slice))) != UIR_OK ||
        (s = uir_check_hip(hipMemcpyAsync(d_sl, h_sl, B * sizeof(uirk_slice), hipMemcpyHostToDevice, st),
                           "hipMemcpyAsync")) != UIR_OK)
        goto out;
    const int n_xcd = uirk_plan_num_xcd();
    float best_fp = 3.0e38f, best_bp = 3.0e38f;
    out->best = (uir_tile){UIRK_FP_ANGLES, 32, UIRK_BP_TILE};
    for (int v = 0; v < 3 && s == UIR_OK; ++v) {
        const uirk_tile tl = {k_fp_variants[v].fp_angles, 32, UIRK_BP_TILE, uir_fp_single_for(spec)};
        uirk_fp_vec fv = {0};
        for (uint32_t i = 0; i < B && s == UIR_OK; ++i)
            if (uirk_plan_fp(&fv, (int)i, (int)n, (int)D, axis, h_sin, h_cos, runs, n_runs, tl.fp_angles,
                             UIRK_TILE_FP_WMAX(tl)))
                s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "work table");
        uirk_xcd_part part;
        free(cost);
        cost = (double *)malloc(sizeof(double) * (fv.n ? fv.n : 1));
        if (s == UIR_OK && !cost)
            s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
        if (s == UIR_OK) {
            for (size_t k = 0; k < fv.n; ++k)
                cost[k] = uirk_fp_item_cost(&fv.v[k]);
            uirk_plan_xcd(cost, (int)fv.n, n_xcd, &part);
        }
        uir_dfree(d_items);
        d_items = NULL;
        if (s == UIR_OK && (s = uir_dmalloc(&d_items, fv.n * sizeof(uirk_fp_item))) == UIR_OK)
            s = uir_check_hip(hipMemcpyAsync(d_items, fv.v, fv.n * sizeof(uirk_fp_item), hipMemcpyHostToDevice, st),
                              "hipMemcpyAsync");
        if (s == UIR_OK)
            s = time_variant(st, 1, 1, d_sl, d_items, (int)fv.n, &part, sdt, idt, &tl,
                             uirk_plan_grid_xcd(&part, (int)fv.n), d_err, reps, &out->fp_ms[v]);
        if (s == UIR_OK && out->fp_ms[v] < best_fp) {
            best_fp = out->fp_ms[v];
            out->best.fp_angles = tl.fp_angles;
        }
        free(fv.v);
    }
    for (int v = 0; v < 4 && s == UIR_OK; ++v) {
        const uirk_tile tl = {UIRK_FP_ANGLES, k_bp_variants[v].bp_angles, k_bp_variants[v].bp_rows, 0};
        uirk_bp_vec bv = {0};
        /* slice pairs where a batch would use them: the paired table of the
         * probe fills the device */
        const size_t tiles = (size_t)((n + UIRK_BP_TILE - 1) / UIRK_BP_TILE) * ((n + tl.bp_rows - 1) / tl.bp_rows);
        const int pairs = B > 1 && (B + 1) / 2 * tiles >= (size_t)uirk_plan_pair_min_items();
        for (uint32_t i = 0; i < B && s == UIR_OK; i += pairs ? 2 : 1) {
            const size_t b0 = bv.n;
            if (uirk_plan_bp(&bv, (int)i, (int)n, tl.bp_rows))
                s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "work table");
            for (size_t k = b0; k < bv.n && pairs && i + 1 < B; ++k)
                bv.v[k].pair = 1;
        }
        uirk_xcd_part part;
        uirk_plan_xcd(NULL, (int)bv.n, n_xcd, &part);
        uir_dfree(d_items);
        d_items = NULL;
        if (s == UIR_OK && (s = uir_dmalloc(&d_items, bv.n * sizeof(uirk_bp_item))) == UIR_OK)
            s = uir_check_hip(hipMemcpyAsync(d_items, bv.v, bv.n * sizeof(uirk_bp_item), hipMemcpyHostToDevice, st),
                              "hipMemcpyAsync");
        if (s == UIR_OK)
            s = time_variant(st, 0, pairs ? 3 : 1, d_sl, d_items, (int)bv.n, &part, sdt, idt, &tl,
                             uirk_plan_grid_xcd(&part, (int)bv.n), d_err, reps, &out->bp_ms[v]);
        if (s == UIR_OK && out->bp_ms[v] < best_bp) {
            best_bp = out->bp_ms[v];
            out->best.bp_angles = tl.bp_angles;
            out->best.bp_rows = tl.bp_rows;
        }
        free(bv.v);
    }
    if (s == UIR_OK) {
        int err = 0;
        s = uir_check_hip(hipMemcpyAsync(&err, d_err, sizeof err, hipMemcpyDeviceToHost, st), "flag read");
        if (s == UIR_OK)
            s = uir_check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
        if (s == UIR_OK && err)
            s = uir_set_error(UIR_ERR_DEVICE, "projector footprint exceeded the staging window (flag %d)", err);
    }
    if (s == UIR_OK)
        s = uir_tune_set(spec, storage, &out->best);
out:
    if (stream)
        (void)hipStreamSynchronize(st);
    uir_dfree(d_items);
    uir_dfree(d_sl);
    uir_dfree(d_err);
    uir_dfree(mem);
    free(cost);
    free(h_sin);
    free(h_cos);
    free(trig);
    free(runs);
    free(h_sl);
    return s;
}

uir_status uir_tune_geometry(uir_resources *res, const uir_slice_spec *spec, uir_dtype storage, uint32_t reps,
                             uir_tune_result *out)
{
    if (!res)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL resources");
    if (spec && spec->model != 0) /* the line model runs the default tile */
        return uir_set_error(UIR_ERR_UNSUPPORTED, "tile tuning is for the joseph model");
    uir_status s = uir_check_hip(hipSetDevice(res->device), "hipSetDevice");
    if (s != UIR_OK)
        return s;
    return uir_tune_run(res->stream, spec, storage, reps, out);
}
