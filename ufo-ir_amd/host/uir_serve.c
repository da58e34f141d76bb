/*
 * uir_serve.c -- reconstruction server with continuous batching.
 *
 * Clients submit requests (geometry, method, iteration counts, parameters and
 * an FP32 sinogram) from any thread.  One worker thread owns the device and a
 * blocking stream.  Its loop:
 *   1. admit queued requests in arrival order while the device budget and the
 *      slot limit allow (a request that does not fit waits at the head of the
 *      queue, so large requests are not starved by small ones);
 *   2. for each admitted request allocate its buffers (one allocation), upload
 *      its sinogram into the storage dtype, and attach its geometry (LUTs,
 *      (sin, cos) table, ray weights R = 1/FP(1) and work-table template),
 *      shared with every other request of the same geometry;
 *   3. when the set of admitted requests changed, rebuild the slice
 *      descriptors and work tables (the geometry templates re-indexed by slot)
 *      and their XCD partitions;
 *   4. run one step: one FP launch and one BP launch over all admitted
 *      requests (plus the split-Bregman TV kernels of SBTV_SIRT requests);
 *   5. retire requests whose iterations are done: copy the image to the
 *      host, release its memory, wake its waiter.
 * At most two steps are in flight, so arrivals are seen within about two step
 * times.  Requests of the line projector model (uir_slice_spec.model = 1)
 * carry its per-angle constants and pixel weights C = 1/BP(1) in their
 * geometry and share the steps in work tables of their own (the table
 * kernels' line mode).  Every FP and BP work item writes only its own slice
 * (two adjacent slices for a BP slice pair, with the arithmetic of one), so
 * a request's result does not depend on what shares its steps.
 *
 * SBTV_SIRT is split-Bregman TV (the outer loop of ufo-ir-sbtv-task.c:
 * b = mu A^T f + lambda D^T (d - b), u-subproblem, shrink, Bregman update)
 * whose u-subproblem  min_u mu/2 |A u - f|^2_R + lambda/2 |d - b - D u|^2  is
 * approximated by `inner_iterations` preconditioned gradient sweeps
 *   u <- u + w P (mu A^T R (f - A u) + lambda D^T (d - b - D u)),
 *   P = 1 / (mu A_n + 8 lambda),
 * A_n = the angle count (= A^T R A 1 for the fast projector pair) and 8 the
 * row-sum bound of D^T D, so P^-1 majorises the subproblem's Hessian and any
 * relaxation w in (0, 1] converges.  u, d and b start at 0.  For the line
 * model A_n is the pixel's BP(1) (= A^T R A 1 up to the rays R drops), so P
 * is an image: the BP kernel forms it per pixel from the weights 1 / BP(1).
 */
#define _GNU_SOURCE
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "uir_internal.h"

enum { RQ_PENDING = 0, RQ_ACTIVE = 1, RQ_DONE = 2, RQ_FAILED = 3, RQ_CANCELLED = 4 };

#define GEOM_CACHE_IDLE 16   /* unreferenced geometries kept for reuse */
#define REQ_SLACK (1u << 20) /* per-request allowance for its work-table entries */

typedef struct sgeom {
    uint32_t n, D, A;
    float step, axis, r_min;
    float *h_sin, *h_cos;
    uir_subset *runs;
    uint32_t n_runs;
    void *mem;          /* ray weights + trig */
    float *d_ray_w, *d_trig;
    uirk_fp_item *fp;   /* template items (slice 0) */
    double *fp_cost;
    size_t n_fp;
    uirk_bp_item *bp;
    size_t n_bp;
    uirk_bp_item *bp32; /* tiles of 32 rows for steps whose BP table would not
                         * fill the device (NULL: the tile has 32 already) */
    size_t n_bp32;
    uint64_t bytes;
    int refs;
    double last_used;
    uirk_tile tile;        /* tuned or default tile shape */
    int fp_shape, bp_shape; /* group indices of that shape */
    int32_t model;          /* 0 joseph, 1 line */
    /* line model: per-angle constants of the table kernels (8 floats) and of
     * the single-angle kernels, and the pixel weights 1 / BP(1) (n x n),
     * carved from mem */
    float *d_lk, *d_pix_w;
    uirk_line_angle *d_line;
    struct sgeom *next;
} sgeom;

/* work-table groups: FP angles 16, 8, 32 with pair staging, then with single
 * values, then the line model; BP 32 or 16 angles x 64 or 32 rows, then the
 * line model */
#define SRV_FP_SHAPES 7
#define SRV_BP_SHAPES 5
#define SRV_FP_LINE 6
#define SRV_BP_LINE 4

typedef struct {
    uirk_tile tile;
    void *h_items, *d_items;
    size_t hcap, dcap;
    int n, grid;
    int mode; /* BP: 1, 3 (slice pairs), 2 (line) or 4 (line pairs) */
    uirk_xcd_part part;
} sgroup;

typedef struct sreq {
    uint64_t ticket;
    uir_request rq;
    float *h_sino;
    float *h_image;
    int state;
    uir_status status;
    char msg[256];
    uir_request_stats stats;
    uint64_t bytes;
    sgeom *g;
    void *mem;
    void *x, *b, *r;
    float *bx, *by, *ex, *ey, *tv;
    uint32_t done_steps, total_steps, inner;
    struct sreq *next_pending;
    struct sreq *next_all;
} sreq;

struct uir_server {
    int device;
    uir_dtype dtype;  /* storage mode */
    int sdt, idt;     /* sinogram and image element types */
    size_t ess, esi;
    uint32_t max_active;
    uint64_t budget, used;
    hipStream_t stream;
    hipEvent_t ev[2];
    int *d_err;
    float *d_tmp;
    size_t tmp_cap;

    pthread_t thread;
    int thread_started;
    pthread_mutex_t mu;
    pthread_cond_t cv_work, cv_done;
    int stop;
    uir_status fatal;
    char fatal_msg[256];

    sreq *pend_head, *pend_tail;
    uint32_t n_pending;
    sreq *all;
    uint64_t next_ticket;
    sreq **active;           /* worker thread only */
    uint32_t n_active, cap_active;
    uint32_t n_live;          /* admitted and not finished (under mu) */
    sgeom *geoms;

    /* tables (worker thread only): slice descriptors, and one work table per
     * tile shape (FP: 6 shapes and the line model's, BP: 4 and the line
     * model's), each launched when non-empty */
    uirk_slice *h_fp_sl, *h_bp_sl;
    double *h_cost;
    uint32_t *ord; /* descriptor slot -> active index */
    size_t hcap_fsl, hcap_bsl, hcap_cost, hcap_ord;
    uirk_slice *d_fp_sl, *d_bp_sl;
    size_t dcap_fsl, dcap_bsl;
    sgroup fpg[SRV_FP_SHAPES], bpg[SRV_BP_SHAPES];
    int autotune;

    uint64_t steps, completed;
    double busy_s, slice_its;
};

#define TRY(x) do { uir_status s_ = (x); if (s_ != UIR_OK) return s_; } while (0)

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

double uir_clock(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- requests ---- */

static int is_sb(const uir_request *r) { return r->method == UIR_METHOD_SBTV_SIRT; }

static uint64_t req_bytes(const uir_request *r, size_t ess, size_t esi)
{
    const uint64_t nv = (uint64_t)r->slice.n * r->slice.n, ns = (uint64_t)r->slice.n_angles * r->slice.n_dets;
    uint64_t b = align_up(nv * esi) + 2 * align_up(ns * ess);
    if (is_sb(r))
        b += 5 * align_up(nv * 4);
    return b + REQ_SLACK;
}

static uint64_t geom_bytes(const uir_slice_spec *sp)
{
    const uint64_t A = sp->n_angles;
    uint64_t b = align_up((size_t)A * sp->n_dets * 4) + align_up((size_t)A * 8);
    if (sp->model == 1)
        b += align_up((size_t)A * 32) + align_up((size_t)sp->n * sp->n * 4) +
             align_up((size_t)A * sizeof(uirk_line_angle));
    return b;
}

static uir_status check_request(const uir_request *r);

uir_status uir_request_footprint(const uir_request *req, uir_dtype storage, uint64_t *bytes)
{
    if (!req || !bytes)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    TRY(check_request(req));
    int sdt = 0, idt = 0;
    if (!uir_storage_split(storage, &sdt, &idt))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "unknown storage dtype");
    *bytes = req_bytes(req, uir_dtype_size(sdt), uir_dtype_size(idt)) +
             geom_bytes(&req->slice);
    return UIR_OK;
}

static uir_status check_request(const uir_request *r)
{
    const uir_slice_spec *s = &r->slice;
    if (s->n == 0 || s->n_dets == 0 || s->n_angles == 0 || s->n > 16384 || s->n_dets > 32768 || s->n_angles > 65536)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "bad geometry %ux%u, %u angles", s->n, s->n_dets, s->n_angles);
    if ((uint64_t)s->n * s->n * 4 >= (1ull << 32) || (uint64_t)s->n_angles * s->n_dets * 4 >= (1ull << 32))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "image or sinogram of 4 GiB or more");
    if (!(s->step >= 0.0f) || !(s->axis_position >= -1.0f))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "bad step or axis");
    if (!uir_geom_active_bins(s->n_dets, uir_geom_axis(s->axis_position, s->n_dets), NULL, NULL))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT,
                             "axis %g leaves no active detector (needs 0.5 <= axis <= n_dets - 0.5)",
                             (double)s->axis_position);
    if (!(s->ray_weight_min >= 0.0f) || !isfinite(s->ray_weight_min))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "bad ray_weight_min");
    if (s->model != 0 && s->model != 1)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "unknown model %d", s->model);
    if (r->iterations == 0)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "iterations must be >= 1");
    if (r->method == UIR_METHOD_SIRT) {
        if (!(s->relaxation >= 0.0f && s->relaxation <= 1.0f))
            return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "relaxation outside [0, 1]");
        return UIR_OK;
    }
    if (r->method != UIR_METHOD_SBTV_SIRT)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "unknown method %u", r->method);
    if (!(s->relaxation > 0.0f && s->relaxation <= 1.0f))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "relaxation outside (0, 1]");
    if (r->inner_iterations == 0)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "inner_iterations must be >= 1");
    if ((uint64_t)r->iterations * r->inner_iterations > 0xffffffffull)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "too many iterations");
    if (!(r->lambda > 0.0f && isfinite(r->lambda)) || !(r->mu > 0.0f && isfinite(r->mu)))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "lambda and mu must be > 0");
    return UIR_OK;
}

static void req_fail(sreq *q, uir_status st, const char *msg)
{
    q->state = st == UIR_ERR_CANCELLED ? RQ_CANCELLED : RQ_FAILED;
    q->status = st;
    snprintf(q->msg, sizeof q->msg, "%s", msg ? msg : "");
}

/* ---- geometries ---- */

static void geom_free(sgeom *g)
{
    if (!g)
        return;
    uir_dfree(g->mem);
    free(g->h_sin);
    free(g->h_cos);
    free(g->runs);
    free(g->fp);
    free(g->fp_cost);
    free(g->bp);
    free(g->bp32);
    free(g);
}

static sgeom *geom_find(uir_server *s, const uir_slice_spec *sp)
{
    const float axis = uir_geom_axis(sp->axis_position, sp->n_dets);
    for (sgeom *g = s->geoms; g; g = g->next)
        if (g->n == sp->n && g->D == sp->n_dets && g->A == sp->n_angles && g->step == sp->step && g->axis == axis &&
            g->r_min == sp->ray_weight_min && g->model == sp->model)
            return g;
    return NULL;
}

/* Creates the geometry of spec (worker thread; the caller has reserved its bytes). */
static uir_status geom_create(uir_server *s, const uir_slice_spec *sp, sgeom **out)
{
    sgeom *g = (sgeom *)calloc(1, sizeof *g);
    if (!g)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
    g->n = sp->n;
    g->D = sp->n_dets;
    g->A = sp->n_angles;
    g->step = sp->step;
    g->axis = uir_geom_axis(sp->axis_position, sp->n_dets);
    g->r_min = sp->ray_weight_min;
    g->model = sp->model;
    g->bytes = geom_bytes(sp);
    g->h_sin = (float *)malloc(sizeof(float) * g->A);
    g->h_cos = (float *)malloc(sizeof(float) * g->A);
    g->runs = (uir_subset *)malloc(sizeof(uir_subset) * g->A);
    float *trig = (float *)malloc(sizeof(float) * 2 * g->A);
    float *h_lk = g->model ? (float *)malloc(sizeof(float) * 8 * g->A) : NULL;
    uirk_line_angle *h_line = g->model ? (uirk_line_angle *)malloc(sizeof(uirk_line_angle) * g->A) : NULL;
    uir_status st = UIR_OK;
    if (!g->h_sin || !g->h_cos || !g->runs || !trig || (g->model && (!h_lk || !h_line))) {
        st = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
        goto fail;
    }
    uir_geom_lut(g->A, g->step, g->h_sin, g->h_cos);
    if (g->model) { /* the unit directions replace the LUT; runs of one family */
        if ((st = uir_line_geom(g->A, g->n, g->axis, g->h_sin, g->h_cos, h_line, h_lk, g->runs, &g->n_runs)) !=
            UIR_OK)
            goto fail;
    } else {
        g->n_runs = uir_geom_runs(g->h_sin, g->h_cos, g->A, g->runs);
    }
    uirk_plan_trig(g->h_sin, g->h_cos, g->A, trig);
    if (s->autotune && !g->model) {
        uir_tile t;
        if (uir_tune_get(sp, s->dtype, &t) == UIR_ERR_NOT_FOUND) {
            uir_tune_result tr;
            if ((st = uir_tune_run(s->stream, sp, s->dtype, 2, &tr)) != UIR_OK)
                goto fail;
        }
    }
    if (g->model) { /* the line model's groups take the default tile */
        g->tile = (uirk_tile)UIRK_TILE_DEFAULT;
        g->fp_shape = SRV_FP_LINE;
        g->bp_shape = SRV_BP_LINE;
    } else {
        g->tile = uir_tile_for(sp, s->dtype);
        g->fp_shape = (g->tile.fp_angles == 8 ? 1 : (g->tile.fp_angles == 32 ? 2 : 0)) + 3 * (g->tile.fp_single != 0);
        g->bp_shape = (g->tile.bp_angles == 16) * 2 + (g->tile.bp_rows == 32);
    }
    {
        uirk_fp_vec fv = {0};
        uirk_bp_vec bv = {0};
        const int fail_fp = g->model ? uirk_plan_fp_line(&fv, 0, (int)g->n, (int)g->D, g->axis, g->h_sin, g->h_cos,
                                                         g->runs, g->n_runs, g->tile.fp_angles,
                                                         UIRK_TILE_FP_WMAX(g->tile))
                                     : uirk_plan_fp(&fv, 0, (int)g->n, (int)g->D, g->axis, g->h_sin, g->h_cos, g->runs,
                                                    g->n_runs, g->tile.fp_angles, UIRK_TILE_FP_WMAX(g->tile));
        uirk_bp_vec bv32 = {0};
        if (fail_fp || uirk_plan_bp(&bv, 0, (int)g->n, g->tile.bp_rows) ||
            (g->tile.bp_rows != 32 && uirk_plan_bp(&bv32, 0, (int)g->n, 32))) {
            free(fv.v);
            free(bv.v);
            free(bv32.v);
            st = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "work table");
            goto fail;
        }
        g->fp = fv.v;
        g->n_fp = fv.n;
        g->bp = bv.v;
        g->n_bp = bv.n;
        g->bp32 = bv32.v;
        g->n_bp32 = bv32.n;
        g->fp_cost = (double *)malloc(sizeof(double) * (g->n_fp ? g->n_fp : 1));
        if (!g->fp_cost) {
            st = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
            goto fail;
        }
        for (size_t i = 0; i < g->n_fp; ++i)
            g->fp_cost[i] = uirk_fp_item_cost(&g->fp[i]);
    }
    if ((st = uir_dmalloc(&g->mem, g->bytes)) != UIR_OK)
        goto fail;
    g->d_ray_w = (float *)g->mem;
    g->d_trig = (float *)((char *)g->mem + align_up((size_t)g->A * g->D * 4));
    if ((st = uir_check_hip(hipMemcpyAsync(g->d_trig, trig, sizeof(float) * 2 * g->A, hipMemcpyHostToDevice,
                                           s->stream), "hipMemcpyAsync")) != UIR_OK)
        goto fail;
    if (g->model) {
        char *p = (char *)g->d_trig + align_up((size_t)g->A * 8);
        g->d_lk = (float *)p;
        p += align_up((size_t)g->A * 32);
        g->d_pix_w = (float *)p;
        p += align_up((size_t)g->n * g->n * 4);
        g->d_line = (uirk_line_angle *)p;
        if ((st = uir_check_hip(hipMemcpyAsync(g->d_lk, h_lk, sizeof(float) * 8 * g->A, hipMemcpyHostToDevice,
                                               s->stream), "hipMemcpyAsync")) != UIR_OK ||
            (st = uir_check_hip(hipMemcpyAsync(g->d_line, h_line, sizeof(uirk_line_angle) * g->A,
                                               hipMemcpyHostToDevice, s->stream), "hipMemcpyAsync")) != UIR_OK ||
            (st = uir_line_weights(s->stream, g->n, g->D, g->A, g->r_min, g->runs, g->n_runs, g->d_line, g->d_ray_w,
                                   g->d_pix_w)) != UIR_OK)
            goto fail;
    } else if ((st = uir_ray_weights(s->stream, g->n, g->D, g->A, g->axis, g->d_trig, g->r_min, g->d_ray_w)) !=
               UIR_OK) {
        goto fail;
    }
    free(trig);
    free(h_lk);
    free(h_line);
    *out = g;
    return UIR_OK;
fail:
    /* the stream may still read the host tables: wait before freeing them */
    (void)hipStreamSynchronize(s->stream);
    free(trig);
    free(h_lk);
    free(h_line);
    geom_free(g);
    return st;
}

/* Frees the least recently used idle geometry; returns its bytes (0: none). */
static uint64_t geom_evict_one(uir_server *s)
{
    sgeom **best = NULL;
    for (sgeom **pp = &s->geoms; *pp; pp = &(*pp)->next)
        if ((*pp)->refs == 0 && (!best || (*pp)->last_used < (*best)->last_used))
            best = pp;
    if (!best)
        return 0;
    sgeom *g = *best;
    *best = g->next;
    const uint64_t b = g->bytes;
    geom_free(g);
    return b;
}

static void geom_trim_idle(uir_server *s)
{
    for (;;) {
        int idle = 0;
        for (sgeom *g = s->geoms; g; g = g->next)
            idle += g->refs == 0;
        if (idle <= GEOM_CACHE_IDLE)
            return;
        pthread_mutex_lock(&s->mu);
        s->used -= geom_evict_one(s);
        pthread_mutex_unlock(&s->mu);
    }
}

/* ---- worker: admission ---- */

/* Bytes admitting q needs now (its buffers, plus its geometry unless resident). */
static uint64_t admit_bytes(uir_server *s, const sreq *q)
{
    return q->bytes + (geom_find(s, &q->rq.slice) ? 0 : geom_bytes(&q->rq.slice));
}

/* Under the lock: pops the requests that fit, in arrival order. */
static uint32_t admit_locked(uir_server *s, sreq **out, uint32_t max_out)
{
    uint32_t k = 0;
    while (s->pend_head && k < max_out) {
        if (s->max_active && s->n_active + k >= s->max_active)
            break;
        sreq *q = s->pend_head;
        uint64_t need = admit_bytes(s, q);
        while (s->used + need > s->budget) {
            const uint64_t freed = geom_evict_one(s);
            if (!freed)
                break;
            s->used -= freed;
            need = admit_bytes(s, q);
        }
        if (s->used + need > s->budget)
            break;
        s->used += need;
        /* pin a resident geometry now, so no later eviction in this pass frees it */
        sgeom *g = geom_find(s, &q->rq.slice);
        if (g) {
            g->refs++;
            q->g = g;
        }
        s->pend_head = q->next_pending;
        if (!s->pend_head)
            s->pend_tail = NULL;
        --s->n_pending;
        q->next_pending = NULL;
        q->state = RQ_ACTIVE;
        q->stats.t_admit = uir_clock();
        s->n_live++;
        out[k++] = q;
    }
    return k;
}

static uir_status tmp_reserve(uir_server *s, size_t elems)
{
    if (elems <= s->tmp_cap)
        return UIR_OK;
    uir_dfree(s->d_tmp);
    s->d_tmp = NULL;
    s->tmp_cap = 0;
    TRY(uir_dmalloc((void **)&s->d_tmp, elems * 4));
    s->tmp_cap = elems;
    return UIR_OK;
}

/* Worker thread: allocate, upload and attach the geometry of an admitted request. */
static uir_status req_setup(uir_server *s, sreq *q)
{
    const uir_slice_spec *sp = &q->rq.slice;
    const size_t nv = (size_t)sp->n * sp->n, ns = (size_t)sp->n_angles * sp->n_dets;
    if (!q->g) {
        sgeom *g = geom_find(s, sp); /* created by an earlier request of this admission pass */
        if (!g) {
            TRY(geom_create(s, sp, &g));
            g->next = s->geoms;
            s->geoms = g;
        }
        g->refs++;
        q->g = g;
    }
    q->g->last_used = uir_clock();
    size_t total = align_up(nv * s->esi) + 2 * align_up(ns * s->ess);
    if (is_sb(&q->rq))
        total += 5 * align_up(nv * 4);
    TRY(uir_dmalloc(&q->mem, total));
    char *p = (char *)q->mem;
    q->x = p;
    p += align_up(nv * s->esi);
    q->b = p;
    p += align_up(ns * s->ess);
    q->r = p;
    p += align_up(ns * s->ess);
    if (is_sb(&q->rq)) {
        float **arr[5] = {&q->bx, &q->by, &q->ex, &q->ey, &q->tv};
        for (int i = 0; i < 5; ++i) {
            *arr[i] = (float *)p;
            p += align_up(nv * 4);
        }
    }
    /* x = 0 and the SB state = 0 (a fill kernel: no memset nodes on this stream) */
    TRY(uir_check_hip(uirk_set((float *)q->mem, total / 4, 0.0f, s->stream), "set"));
    TRY(tmp_reserve(s, ns > nv ? ns : nv));
    TRY(uir_check_hip(hipMemcpyAsync(s->d_tmp, q->h_sino, ns * 4, hipMemcpyHostToDevice, s->stream), "upload"));
    TRY(uir_check_hip(uirk_convert(s->d_tmp, UIR_DTYPE_F32, q->b, s->sdt, ns, s->stream), "convert"));
    TRY(uir_check_hip(hipStreamSynchronize(s->stream), "hipStreamSynchronize"));
    free(q->h_sino);
    q->h_sino = NULL;
    q->total_steps = is_sb(&q->rq) ? q->rq.iterations * q->rq.inner_iterations : q->rq.iterations;
    q->done_steps = 0;
    q->inner = 0;
    return UIR_OK;
}

/* Releases an active request's device memory and geometry reference. */
static void req_release(uir_server *s, sreq *q)
{
    uir_dfree(q->mem);
    q->mem = NULL;
    if (q->g) {
        q->g->refs--;
        q->g->last_used = uir_clock();
        q->g = NULL;
    }
    pthread_mutex_lock(&s->mu);
    s->used -= q->bytes;
    pthread_mutex_unlock(&s->mu);
}

/* ---- worker: tables ---- */

static int grow(void **p, size_t *cap, size_t need, size_t es)
{
    if (need <= *cap)
        return 0;
    size_t c = *cap ? *cap : 64;
    while (c < need)
        c *= 2;
    void *np = realloc(*p, c * es);
    if (!np)
        return -1;
    *p = np;
    *cap = c;
    return 0;
}

static uir_status dgrow(void **p, size_t *cap, size_t need, size_t es)
{
    if (need <= *cap)
        return UIR_OK;
    size_t c = *cap ? *cap : 64;
    while (c < need)
        c *= 2;
    uir_dfree(*p);
    *p = NULL;
    *cap = 0;
    TRY(uir_dmalloc(p, c * es));
    *cap = c;
    return UIR_OK;
}

/* Rebuilds descriptors and work tables of the admitted set.  The stream is
 * idle (the caller synchronised it), so the device tables can be rewritten.
 * Descriptors are laid out geometry by geometry (slot j holds request
 * ord[j]), so that consecutive requests of one geometry share BP items as
 * slice pairs (modes 3 and 4). */
/* A geometry's BP template items and group in a step: its tuned tile, or
 * tiles of 32 rows when the step's table would not fill the device (rows32) */
static const uirk_bp_item *geom_bp(const sgeom *g, int rows32, size_t *n)
{
    const int r32 = rows32 && g->bp32;
    *n = r32 ? g->n_bp32 : g->n_bp;
    return r32 ? g->bp32 : g->bp;
}
static int geom_bp_shape(const sgeom *g, int rows32)
{
    return g->bp_shape == SRV_BP_LINE ? SRV_BP_LINE : (rows32 ? g->bp_shape | 1 : g->bp_shape);
}

static uir_status tables_build(uir_server *s)
{
    const uint32_t na = s->n_active;
    size_t nf[SRV_FP_SHAPES] = {0}, nb[SRV_BP_SHAPES] = {0}, ncost = 1, bp_items = 0;
    for (uint32_t i = 0; i < na; ++i)
        bp_items += s->active[i]->g->n_bp;
    const int rows32 = uirk_plan_bp_rows32(bp_items);
    s->bpg[SRV_BP_LINE].tile.bp_rows = rows32 ? 32 : UIRK_BP_TILE;
    for (uint32_t i = 0; i < na; ++i) {
        const sgeom *g = s->active[i]->g;
        size_t n;
        (void)geom_bp(g, rows32, &n);
        nf[g->fp_shape] += g->n_fp;
        nb[geom_bp_shape(g, rows32)] += n;
    }
    for (int c = 0; c < SRV_FP_SHAPES; ++c)
        ncost = nf[c] > ncost ? nf[c] : ncost;
    for (int c = 0; c < SRV_BP_SHAPES; ++c)
        ncost = nb[c] > ncost ? nb[c] : ncost;
    if (grow((void **)&s->h_fp_sl, &s->hcap_fsl, na, sizeof(uirk_slice)) ||
        grow((void **)&s->h_bp_sl, &s->hcap_bsl, na, sizeof(uirk_slice)) ||
        grow((void **)&s->h_cost, &s->hcap_cost, ncost, sizeof(double)) ||
        grow((void **)&s->ord, &s->hcap_ord, na, sizeof(uint32_t)))
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "work table");
    {
        /* ord: active requests grouped by geometry, in order of first arrival */
        uint32_t k = 0;
        for (uint32_t i = 0; i < na; ++i) {
            int seen = 0;
            for (uint32_t j = 0; j < i && !seen; ++j)
                seen = s->active[j]->g == s->active[i]->g;
            if (seen)
                continue;
            for (uint32_t j = i; j < na; ++j)
                if (s->active[j]->g == s->active[i]->g)
                    s->ord[k++] = j;
        }
    }
    for (int c = 0; c < SRV_FP_SHAPES; ++c)
        if (grow(&s->fpg[c].h_items, &s->fpg[c].hcap, nf[c], sizeof(uirk_fp_item)))
            return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "work table");
    for (int c = 0; c < SRV_BP_SHAPES; ++c)
        if (grow(&s->bpg[c].h_items, &s->bpg[c].hcap, nb[c], sizeof(uirk_bp_item)))
            return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "work table");
    for (int c = 0; c < SRV_FP_SHAPES; ++c)
        s->fpg[c].n = 0;
    for (int c = 0; c < SRV_BP_SHAPES; ++c)
        s->bpg[c].n = 0;
    /* BP pairs only while the paired table of the model fills the device */
    size_t paired[2] = {0, 0};
    for (uint32_t i = 0; i < na; ++i) {
        const sgeom *g = s->active[s->ord[i]]->g;
        size_t n;
        (void)geom_bp(g, rows32, &n);
        paired[g->model != 0] += n;
        if (i + 1 < na && s->active[s->ord[i + 1]]->g == g)
            ++i;
    }
    const int pair_ok[2] = {paired[0] >= (size_t)uirk_plan_pair_min_items(),
                            paired[1] >= (size_t)uirk_plan_pair_min_items()};
    int lead = 0; /* slot i - 1 led a BP pair */
    for (uint32_t i = 0; i < na; ++i) {
        const sreq *q = s->active[s->ord[i]];
        const sgeom *g = q->g;
        uirk_slice *f = &s->h_fp_sl[i], *k = &s->h_bp_sl[i];
        memset(f, 0, sizeof *f);
        f->x = q->x;
        f->b = q->b;
        f->r = q->r;
        f->ray_w = g->d_ray_w;
        f->trig = g->d_trig;
        f->n = (int32_t)g->n;
        f->n_dets = (int32_t)g->D;
        f->n_angles = (int32_t)g->A;
        f->axis = g->axis;
        f->corr = -1.0f;
        f->n_runs = 1;
        f->run_off[0] = 0;
        f->run_n[0] = g->A;
        *k = *f;
        k->x_out = q->x;
        k->b = q->r;
        k->r = NULL;
        k->ray_w = NULL;
        k->posc = q->rq.slice.positive_constraint ? 1 : 0;
        k->relax = q->rq.slice.relaxation;
        if (g->model) { /* line: its constants, and per-pixel weights in place of 1/A */
            f->line = g->d_lk;
            k->line = g->d_lk;
            k->pix_w = g->d_pix_w;
            k->pix_scale = 1.0f;
            if (is_sb(&q->rq)) { /* the kernel forms P per pixel from pix_w */
                k->pix_scale = q->rq.mu;
                k->tv = q->tv;
                k->tv_scale = q->rq.lambda;
            }
        } else if (is_sb(&q->rq)) {
            const float P = 1.0f / (q->rq.mu * (float)g->A + 8.0f * q->rq.lambda);
            k->pix_scale = q->rq.mu * P;
            k->tv = q->tv;
            k->tv_scale = q->rq.slice.relaxation * q->rq.lambda * P;
        } else {
            k->pix_scale = 1.0f / (float)g->A;
        }
        sgroup *fg = &s->fpg[g->fp_shape], *bg = &s->bpg[geom_bp_shape(g, rows32)];
        uirk_fp_item *fi = (uirk_fp_item *)fg->h_items;
        uirk_bp_item *bi = (uirk_bp_item *)bg->h_items;
        size_t n_bp;
        const uirk_bp_item *gbp = geom_bp(g, rows32, &n_bp);
        for (size_t j = 0; j < g->n_fp; ++j) {
            fi[fg->n] = g->fp[j];
            fi[fg->n].slice = (int32_t)i;
            ++fg->n;
        }
        /* a slot followed by one of its geometry leads a BP pair (when the
         * model's paired table fills the device); the second slot of a pair
         * has no BP items of its own */
        const int second = lead;
        lead = pair_ok[g->model != 0] && !second && i + 1 < na && s->active[s->ord[i + 1]]->g == g;
        for (size_t j = 0; j < n_bp && !second; ++j) {
            bi[bg->n] = gbp[j];
            bi[bg->n].slice = (int32_t)i;
            bi[bg->n].pair = lead;
            ++bg->n;
        }
    }
    /* per group: cost-balanced XCD ranges (the items stay slice-major) */
    for (int c = 0; c < SRV_FP_SHAPES; ++c) {
        sgroup *grp = &s->fpg[c];
        const uirk_fp_item *it = (const uirk_fp_item *)grp->h_items;
        for (int j = 0; j < grp->n; ++j)
            s->h_cost[j] = uirk_fp_item_cost(&it[j]);
        uirk_plan_xcd(s->h_cost, grp->n, uirk_plan_num_xcd(), &grp->part);
        grp->grid = uirk_plan_grid_xcd(&grp->part, (int)grp->n);
    }
    for (int c = 0; c < SRV_BP_SHAPES; ++c) {
        sgroup *grp = &s->bpg[c];
        const uirk_bp_item *it = (const uirk_bp_item *)grp->h_items;
        grp->mode = c == SRV_BP_LINE ? 2 : 1;
        /* an item costs its angles, paired or not: in a launch with pairs,
         * an unpaired item runs the pair kernel and costs what a pair does */
        for (int j = 0; j < grp->n; ++j) {
            s->h_cost[j] = (double)s->h_bp_sl[it[j].slice].n_angles;
            grp->mode = it[j].pair ? (c == SRV_BP_LINE ? 4 : 3) : grp->mode;
        }
        uirk_plan_xcd(s->h_cost, grp->n, uirk_plan_num_xcd(), &grp->part);
        grp->grid = uirk_plan_grid_xcd(&grp->part, (int)grp->n);
    }
    TRY(dgrow((void **)&s->d_fp_sl, &s->dcap_fsl, na, sizeof(uirk_slice)));
    TRY(dgrow((void **)&s->d_bp_sl, &s->dcap_bsl, na, sizeof(uirk_slice)));
    TRY(uir_check_hip(hipMemcpy(s->d_fp_sl, s->h_fp_sl, na * sizeof(uirk_slice), hipMemcpyHostToDevice), "tables"));
    TRY(uir_check_hip(hipMemcpy(s->d_bp_sl, s->h_bp_sl, na * sizeof(uirk_slice), hipMemcpyHostToDevice), "tables"));
    for (int c = 0; c < SRV_FP_SHAPES; ++c) {
        sgroup *grp = &s->fpg[c];
        if (!grp->n)
            continue;
        TRY(dgrow(&grp->d_items, &grp->dcap, (size_t)grp->n, sizeof(uirk_fp_item)));
        TRY(uir_check_hip(hipMemcpy(grp->d_items, grp->h_items, (size_t)grp->n * sizeof(uirk_fp_item),
                                    hipMemcpyHostToDevice), "tables"));
    }
    for (int c = 0; c < SRV_BP_SHAPES; ++c) {
        sgroup *grp = &s->bpg[c];
        if (!grp->n)
            continue;
        TRY(dgrow(&grp->d_items, &grp->dcap, (size_t)grp->n, sizeof(uirk_bp_item)));
        TRY(uir_check_hip(hipMemcpy(grp->d_items, grp->h_items, (size_t)grp->n * sizeof(uirk_bp_item),
                                    hipMemcpyHostToDevice), "tables"));
    }
    return UIR_OK;
}

/* ---- worker: steps ---- */

static uir_status step_launch(uir_server *s)
{
    for (uint32_t i = 0; i < s->n_active; ++i) {
        sreq *q = s->active[i];
        if (is_sb(&q->rq))
            TRY(uir_check_hip(uirk_sb_tv(q->x, s->idt, q->ex, q->ey, q->tv, q->rq.slice.n, s->stream), "sb_tv"));
    }
    for (int c = 0; c < SRV_FP_SHAPES; ++c)
        if (s->fpg[c].n)
            TRY(uir_check_hip(uirk_fp_table(s->d_fp_sl, NULL, (const uirk_fp_item *)s->fpg[c].d_items, s->fpg[c].n,
                                            &s->fpg[c].part, c == SRV_FP_LINE ? 2 : 1, s->sdt, s->idt,
                                            &s->fpg[c].tile, s->fpg[c].grid, s->d_err, s->stream),
                              "FP launch"));
    for (int c = 0; c < SRV_BP_SHAPES; ++c)
        if (s->bpg[c].n)
            TRY(uir_check_hip(uirk_bp_table(s->d_bp_sl, NULL, (const uirk_bp_item *)s->bpg[c].d_items, s->bpg[c].n,
                                            &s->bpg[c].part, s->bpg[c].mode, s->sdt, s->idt, &s->bpg[c].tile,
                                            s->bpg[c].grid, s->d_err, s->stream),
                              "BP launch"));
    for (uint32_t i = 0; i < s->n_active; ++i) {
        sreq *q = s->active[i];
        q->done_steps++;
        q->stats.steps++;
        if (s->n_active > q->stats.max_batch)
            q->stats.max_batch = s->n_active;
        if (is_sb(&q->rq) && ++q->inner == q->rq.inner_iterations) {
            q->inner = 0;
            /* the last outer iteration's Bregman update does not change u */
            if (q->done_steps < q->total_steps)
                TRY(uir_check_hip(uirk_sb_shrink(q->x, s->idt, q->bx, q->by, q->ex, q->ey, q->rq.slice.n,
                                                 q->rq.lambda, s->stream), "sb_shrink"));
        }
    }
    return UIR_OK;
}

/* Copies a finished request's image to the host (synchronises the stream). */
static uir_status req_finish(uir_server *s, sreq *q)
{
    const size_t nv = (size_t)q->rq.slice.n * q->rq.slice.n;
    int err = 0;
    TRY(uir_check_hip(hipMemcpyAsync(&err, s->d_err, sizeof err, hipMemcpyDeviceToHost, s->stream), "flag read"));
    TRY(tmp_reserve(s, nv));
    TRY(uir_check_hip(uirk_convert(q->x, s->idt, s->d_tmp, UIR_DTYPE_F32, nv, s->stream), "convert"));
    q->h_image = (float *)malloc(nv * 4);
    if (!q->h_image)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
    TRY(uir_check_hip(hipMemcpyAsync(q->h_image, s->d_tmp, nv * 4, hipMemcpyDeviceToHost, s->stream), "download"));
    TRY(uir_check_hip(hipStreamSynchronize(s->stream), "hipStreamSynchronize"));
    if (err) /* the server then fails every request of those steps and stops */
        return uir_set_error(UIR_ERR_DEVICE, "projector footprint exceeded the staging window (flag %d)", err);
    return UIR_OK;
}

static void fail_all_locked(uir_server *s, uir_status st, const char *msg)
{
    s->fatal = st;
    snprintf(s->fatal_msg, sizeof s->fatal_msg, "%s", msg);
    for (uint32_t i = 0; i < s->n_active; ++i)
        req_fail(s->active[i], st, msg);
    for (sreq *q = s->pend_head; q; q = q->next_pending)
        req_fail(q, st, msg);
    s->pend_head = s->pend_tail = NULL;
    s->n_pending = 0;
    pthread_cond_broadcast(&s->cv_done);
}

static void *worker(void *arg)
{
    uir_server *s = (uir_server *)arg;
    (void)hipSetDevice(s->device);
    sreq **adm = NULL;
    uint32_t adm_cap = 0;
    int in_flight = 0;   /* steps enqueued since the last synchronisation (<= 2) */
    uint64_t ev_step = 0;
    double t_busy0 = 0;

    pthread_mutex_lock(&s->mu);
    for (;;) {
        while (!s->stop && !s->pend_head && s->n_active == 0)
            pthread_cond_wait(&s->cv_work, &s->mu);
        if (s->stop && s->n_active == 0 && !s->pend_head)
            break;
        if (s->fatal != UIR_OK) {
            fail_all_locked(s, s->fatal, s->fatal_msg);
            if (s->stop)
                break;
            pthread_cond_wait(&s->cv_work, &s->mu);
            continue;
        }
        /* admission */
        if (adm_cap < s->n_pending + 1) {
            sreq **na = (sreq **)realloc(adm, sizeof(sreq *) * (s->n_pending + 16));
            if (na) {
                adm = na;
                adm_cap = s->n_pending + 16;
            }
        }
        uint32_t n_adm = adm ? admit_locked(s, adm, adm_cap) : 0;
        if (n_adm && s->n_active + n_adm > s->cap_active) {
            uint32_t c = s->cap_active ? s->cap_active : 16;
            while (c < s->n_active + n_adm)
                c *= 2;
            sreq **na = (sreq **)realloc(s->active, sizeof(sreq *) * c);
            if (!na) {
                fail_all_locked(s, UIR_ERR_OUT_OF_MEMORY, "active list");
                continue;
            }
            s->active = na;
            s->cap_active = c;
        }
        pthread_mutex_unlock(&s->mu);

        uir_status st = UIR_OK;
        int changed = 0;
        if (n_adm) {
            if (in_flight) {
                st = uir_check_hip(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
                in_flight = 0;
            }
            for (uint32_t i = 0; i < n_adm && st == UIR_OK; ++i) {
                sreq *q = adm[i];
                uir_status rs = req_setup(s, q);
                if (rs != UIR_OK) {
                    char msg[256];
                    snprintf(msg, sizeof msg, "%s", uir_last_error());
                    req_release(s, q);
                    pthread_mutex_lock(&s->mu);
                    /* its geometry bytes were reserved too if the geometry was missing */
                    s->n_live--;
                    req_fail(q, rs, msg);
                    pthread_cond_broadcast(&s->cv_done);
                    pthread_mutex_unlock(&s->mu);
                    if (rs == UIR_ERR_DEVICE)
                        st = rs;
                    continue;
                }
                if (s->n_active == 0)
                    t_busy0 = uir_clock();
                s->active[s->n_active++] = q;
                changed = 1;
            }
            /* account geometry bytes exactly: used = requests + resident geometries + ... */
            pthread_mutex_lock(&s->mu);
            {
                uint64_t u = 0;
                for (uint32_t i = 0; i < s->n_active; ++i)
                    u += s->active[i]->bytes;
                for (sgeom *g = s->geoms; g; g = g->next)
                    u += g->bytes;
                s->used = u;
            }
            pthread_mutex_unlock(&s->mu);
        }
        if (st == UIR_OK && changed && s->n_active)
            st = tables_build(s);

        /* one step */
        if (st == UIR_OK && s->n_active) {
            if (in_flight == 2) {
                st = uir_check_hip(hipEventSynchronize(s->ev[ev_step & 1]), "hipEventSynchronize");
                in_flight = 1;
            }
            if (st == UIR_OK)
                st = step_launch(s);
            if (st == UIR_OK)
                st = uir_check_hip(hipEventRecord(s->ev[ev_step & 1], s->stream), "hipEventRecord");
            ++ev_step;
            ++in_flight;
            s->steps++;
            s->slice_its += s->n_active;
        }
        /* retire */
        if (st == UIR_OK) {
            uint32_t keep = 0;
            int retired = 0;
            for (uint32_t i = 0; i < s->n_active; ++i) {
                sreq *q = s->active[i];
                if (q->done_steps < q->total_steps) {
                    s->active[keep++] = q;
                    continue;
                }
                /* The device flag is shared by every request of the steps since
                 * the last read: once it is found set, every request retiring in
                 * this pass fails too (the active ones fail with the server). */
                uir_status rs = st == UIR_ERR_DEVICE
                                    ? uir_set_error(UIR_ERR_DEVICE, "projector footprint exceeded the staging window "
                                                                    "in a step this request took part in")
                                    : req_finish(s, q);
                in_flight = 0;
                char msg[256];
                snprintf(msg, sizeof msg, "%s", rs == UIR_OK ? "" : uir_last_error());
                req_release(s, q);
                pthread_mutex_lock(&s->mu);
                s->n_live--;
                if (rs == UIR_OK) {
                    q->state = RQ_DONE;
                    q->status = UIR_OK;
                    q->stats.t_done = uir_clock();
                    s->completed++;
                } else {
                    req_fail(q, rs, msg);
                }
                pthread_cond_broadcast(&s->cv_done);
                pthread_mutex_unlock(&s->mu);
                retired = 1;
                if (rs == UIR_ERR_DEVICE)
                    st = rs;
            }
            if (retired) {
                s->n_active = keep;
                if (keep)
                    st = st == UIR_OK ? tables_build(s) : st;
                else
                    s->busy_s += uir_clock() - t_busy0;
                geom_trim_idle(s);
            }
        }
        if (st != UIR_OK) {
            /* release device memory first: a woken waiter frees its request */
            char msg[256];
            snprintf(msg, sizeof msg, "%s", uir_last_error());
            (void)hipStreamSynchronize(s->stream);
            in_flight = 0;
            for (uint32_t i = 0; i < s->n_active; ++i)
                req_release(s, s->active[i]);
            pthread_mutex_lock(&s->mu);
            fail_all_locked(s, st, msg);
            s->n_active = 0;
            s->n_live = 0;
        } else {
            pthread_mutex_lock(&s->mu);
        }
    }
    pthread_mutex_unlock(&s->mu);
    free(adm);
    return NULL;
}

/* ---- public API ---- */

uir_status uir_server_new(int device, const uir_server_config *cfg, uir_server **out)
{
    if (!cfg || !out)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = NULL;
    int sdt = 0, idt = 0;
    if (!uir_storage_split(cfg->storage, &sdt, &idt))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "unknown storage dtype");
    int nd = 0;
    TRY(uir_device_count(&nd));
    if (device < 0 || device >= nd)
        return uir_set_error(UIR_ERR_NO_DEVICE, "device %d of %d", device, nd);
    TRY(uir_check_hip(hipSetDevice(device), "hipSetDevice"));
    uint64_t budget = cfg->hbm_budget;
    if (!budget) {
        size_t f = 0;
        TRY(uir_mem_get_info(&f, NULL));
        if (f <= cfg->hbm_reserve)
            return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "%zu bytes free, %llu reserved", f,
                                 (unsigned long long)cfg->hbm_reserve);
        budget = f - cfg->hbm_reserve;
    }
    uir_server *s = (uir_server *)calloc(1, sizeof *s);
    if (!s)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
    s->device = device;
    s->dtype = cfg->storage;
    s->sdt = sdt;
    s->idt = idt;
    s->ess = uir_dtype_size(sdt);
    s->esi = uir_dtype_size(idt);
    s->max_active = cfg->max_active;
    s->budget = budget;
    s->next_ticket = 1;
    s->autotune = cfg->autotune;
    for (int c = 0; c < SRV_FP_SHAPES; ++c) {
        s->fpg[c].tile = (uirk_tile)UIRK_TILE_DEFAULT;
        if (c == SRV_FP_LINE)
            continue;
        s->fpg[c].tile.fp_angles = c % 3 == 1 ? 8 : (c % 3 == 2 ? 32 : UIRK_FP_ANGLES);
        s->fpg[c].tile.fp_single = c / 3;
    }
    for (int c = 0; c < SRV_BP_SHAPES; ++c) {
        s->bpg[c].tile = (uirk_tile)UIRK_TILE_DEFAULT;
        if (c == SRV_BP_LINE)
            continue;
        s->bpg[c].tile.bp_angles = c >= 2 ? 16 : 32;
        s->bpg[c].tile.bp_rows = (c & 1) ? 32 : UIRK_BP_TILE;
    }
    s->fatal = UIR_OK;
    pthread_mutex_init(&s->mu, NULL);
    pthread_cond_init(&s->cv_work, NULL);
    pthread_cond_init(&s->cv_done, NULL);
    uir_status st = uir_check_hip(hipStreamCreate(&s->stream), "hipStreamCreate");
    if (st == UIR_OK)
        st = uir_check_hip(hipEventCreateWithFlags(&s->ev[0], hipEventDisableTiming), "hipEventCreate");
    if (st == UIR_OK)
        st = uir_check_hip(hipEventCreateWithFlags(&s->ev[1], hipEventDisableTiming), "hipEventCreate");
    if (st == UIR_OK)
        st = uir_dmalloc((void **)&s->d_err, sizeof(int));
    if (st == UIR_OK)
        st = uir_check_hip(hipMemset(s->d_err, 0, sizeof(int)), "hipMemset");
    if (st == UIR_OK) {
        if (pthread_create(&s->thread, NULL, worker, s) != 0)
            st = uir_set_error(UIR_ERR_DEVICE, "pthread_create failed");
        else
            s->thread_started = 1;
    }
    if (st != UIR_OK) {
        uir_server_free(s);
        return st;
    }
    *out = s;
    return UIR_OK;
}

uir_status uir_server_submit(uir_server *s, const uir_request *req, const float *sinogram, uint64_t *ticket)
{
    if (!s || !req || !sinogram || !ticket)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    TRY(check_request(req));
    const uint64_t need = req_bytes(req, s->ess, s->esi) + geom_bytes(&req->slice);
    if (need > s->budget)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "request needs %llu device bytes, the budget is %llu",
                             (unsigned long long)need, (unsigned long long)s->budget);
    const size_t ns = (size_t)req->slice.n_angles * req->slice.n_dets;
    sreq *q = (sreq *)calloc(1, sizeof *q);
    if (!q)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
    q->h_sino = (float *)malloc(ns * 4);
    if (!q->h_sino) {
        free(q);
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
    }
    memcpy(q->h_sino, sinogram, ns * 4);
    q->rq = *req;
    q->bytes = req_bytes(req, s->ess, s->esi);
    q->stats.device_bytes = q->bytes;
    q->stats.t_submit = uir_clock();
    q->state = RQ_PENDING;
    pthread_mutex_lock(&s->mu);
    if (s->stop || s->fatal != UIR_OK) {
        const uir_status fs = s->fatal != UIR_OK ? s->fatal : UIR_ERR_CANCELLED;
        char msg[256];
        snprintf(msg, sizeof msg, "%s", s->fatal != UIR_OK ? s->fatal_msg : "server is stopping");
        pthread_mutex_unlock(&s->mu);
        free(q->h_sino);
        free(q);
        return uir_set_error(fs, "%s", msg);
    }
    q->ticket = s->next_ticket++;
    q->next_all = s->all;
    s->all = q;
    if (s->pend_tail)
        s->pend_tail->next_pending = q;
    else
        s->pend_head = q;
    s->pend_tail = q;
    s->n_pending++;
    *ticket = q->ticket;
    pthread_cond_signal(&s->cv_work);
    pthread_mutex_unlock(&s->mu);
    return UIR_OK;
}

uir_status uir_server_wait(uir_server *s, uint64_t ticket, float *image, double timeout_s, uir_request_stats *stats)
{
    if (!s)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL server");
    struct timespec dl;
    if (timeout_s >= 0) {
        clock_gettime(CLOCK_REALTIME, &dl);
        double t = (double)dl.tv_sec + 1e-9 * dl.tv_nsec + timeout_s;
        dl.tv_sec = (time_t)t;
        dl.tv_nsec = (long)((t - (double)dl.tv_sec) * 1e9);
    }
    pthread_mutex_lock(&s->mu);
    sreq **pp = &s->all;
    while (*pp && (*pp)->ticket != ticket)
        pp = &(*pp)->next_all;
    if (!*pp) {
        pthread_mutex_unlock(&s->mu);
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
    }
    sreq *q = *pp;
    while (q->state == RQ_PENDING || q->state == RQ_ACTIVE) {
        int rc = timeout_s >= 0 ? pthread_cond_timedwait(&s->cv_done, &s->mu, &dl)
                                : pthread_cond_wait(&s->cv_done, &s->mu);
        if (rc == ETIMEDOUT && (q->state == RQ_PENDING || q->state == RQ_ACTIVE)) {
            pthread_mutex_unlock(&s->mu);
            return uir_set_error(UIR_ERR_TIMEOUT, "request %llu not done", (unsigned long long)ticket);
        }
    }
    /* unlink: the ticket is released */
    pp = &s->all;
    while (*pp != q)
        pp = &(*pp)->next_all;
    *pp = q->next_all;
    pthread_mutex_unlock(&s->mu);
    uir_status st = q->status;
    if (st == UIR_OK) {
        if (image)
            memcpy(image, q->h_image, (size_t)q->rq.slice.n * q->rq.slice.n * 4);
    } else {
        uir_set_error(st, "%s", q->msg);
    }
    if (stats)
        *stats = q->stats;
    free(q->h_image);
    free(q->h_sino);
    free(q);
    return st;
}

uir_status uir_server_get_info(uir_server *s, uir_server_info *info)
{
    if (!s || !info)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    pthread_mutex_lock(&s->mu);
    info->budget = s->budget;
    info->used = s->used;
    info->active = s->n_live;
    info->pending = s->n_pending;
    info->steps = s->steps;
    info->completed = s->completed;
    info->busy_s = s->busy_s;
    info->slice_iterations = s->slice_its;
    pthread_mutex_unlock(&s->mu);
    return UIR_OK;
}

void uir_server_free(uir_server *s)
{
    if (!s)
        return;
    if (s->thread_started) {
        pthread_mutex_lock(&s->mu);
        for (sreq *q = s->pend_head; q; q = q->next_pending)
            req_fail(q, UIR_ERR_CANCELLED, "server freed");
        s->pend_head = s->pend_tail = NULL;
        s->n_pending = 0;
        s->stop = 1;
        pthread_cond_broadcast(&s->cv_work);
        pthread_cond_broadcast(&s->cv_done);
        pthread_mutex_unlock(&s->mu);
        pthread_join(s->thread, NULL);
    }
    if (s->stream)
        (void)hipStreamSynchronize(s->stream);
    for (uint32_t i = 0; i < s->n_active; ++i)
        uir_dfree(s->active[i]->mem);
    while (s->all) {
        sreq *q = s->all;
        s->all = q->next_all;
        free(q->h_image);
        free(q->h_sino);
        free(q);
    }
    while (s->geoms) {
        sgeom *g = s->geoms;
        s->geoms = g->next;
        geom_free(g);
    }
    free(s->active);
    free(s->h_fp_sl);
    free(s->h_bp_sl);
    free(s->h_cost);
    free(s->ord);
    for (int c = 0; c < SRV_FP_SHAPES; ++c) {
        free(s->fpg[c].h_items);
        uir_dfree(s->fpg[c].d_items);
    }
    for (int c = 0; c < SRV_BP_SHAPES; ++c) {
        free(s->bpg[c].h_items);
        uir_dfree(s->bpg[c].d_items);
    }
    uir_dfree(s->d_fp_sl);
    uir_dfree(s->d_bp_sl);

    uir_dfree(s->d_tmp);
    uir_dfree(s->d_err);
    if (s->ev[0])
        (void)hipEventDestroy(s->ev[0]);
    if (s->ev[1])
        (void)hipEventDestroy(s->ev[1]);
    if (s->stream)
        (void)hipStreamDestroy(s->stream);
    pthread_mutex_destroy(&s->mu);
    pthread_cond_destroy(&s->cv_work);
    pthread_cond_destroy(&s->cv_done);
    free(s);
}
