/*
 * uir_asdpocs.c -- batched ASD-POCS over the slices of a uir_batch
 * (ufo-ir-asdpocs-task.c:411-509 with a SART df-minimizer,
 * ufo-ir-sart-task.c:173-228, and tvstd :550-580).
 *
 * One iteration, for every slice at once:
 *   data step   x = 0 (reference) or x_prev = x (textbook); df_iterations SART
 *               passes: per angle k, one FP launch writing the residual rows
 *               r[k] = (b[k] - A_k x) R[k] of every slice that has angle k
 *               (uirk_fp_angle) and one streaming BP launch adding
 *               relax A_k^T r[k] to the images (uirk_bp_angle);
 *   POSC + dp   x = max(x, 0), dp = |x - x_prev|, x_prev = x;
 *   dd          r = b - A x over all angles (the batch's FP tables), |r|;
 *   TV          ng times: g = l1_grad(x), |g|, x -= dtgv / |g| g;
 *   dg          |x - x_prev|, then the dtgv / beta updates.
 * Norms are FP64 reductions on the device (L1 for the reference, L2 for the
 * textbook variant) and every scalar the method needs lives in a per-slice
 * state in device memory, so an iteration has no host synchronisation and is
 * captured once as a hipGraph.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "uir_batch_impl.h"

#define TRY(x) do { uir_status s_ = (x); if (s_ != UIR_OK) return s_; } while (0)
#define HIP(x, what) TRY(uir_check_hip((x), (what)))

struct uir_asd {
    void *mem;                 /* x_prev (image dtype) and g (FP32) of every slice */
    uirk_asd_slice *d_sl;
    uirk_slice *h_bp_sl;       /* SART BP descriptors (relax reset at each call) */
    uirk_slice *d_fp_sl, *d_dd_sl, *d_bp_sl;
    uirk_fp_item *d_items;     /* per-angle FP items: {slice, d0} per 128 detectors */
    int n_items, grid;
    uirk_xcd_part part;
    uint32_t n_max, a_max;
    uirk_asd_state *d_state;
    double *d_partials;
    int nblk;
    float *d_diag;             /* diag_cap rows of (dd, dp, dg, dtgv) per slice */
    uint32_t diag_cap, diag_rows;
    hipGraph_t graph;
    hipGraphExec_t exec;
    uir_asdpocs_params graph_p; /* parameters the graph was captured with */
};

void uir_asd_free(struct uir_asd *a)
{
    if (!a)
        return;
    if (a->exec)
        (void)hipGraphExecDestroy(a->exec);
    if (a->graph)
        (void)hipGraphDestroy(a->graph);
    uir_dfree(a->mem);
    uir_dfree(a->d_sl);
    uir_dfree(a->d_fp_sl);
    uir_dfree(a->d_dd_sl);
    uir_dfree(a->d_bp_sl);
    uir_dfree(a->d_items);
    uir_dfree(a->d_state);
    uir_dfree(a->d_partials);
    uir_dfree(a->d_diag);
    free(a->h_bp_sl);
    free(a);
}

void uir_asdpocs_params_default(uir_asdpocs_params *p)
{
    if (!p)
        return;
    memset(p, 0, sizeof *p);
    p->beta = 1.0f;        /* ufo-ir-asdpocs-task.c:173-179 */
    p->beta_red = 0.995f;
    p->alpha = 0.2f;
    p->alpha_red = 0.95f;
    p->r_max = 0.95f;
    p->ng = 20;
    p->df_iterations = 1;
    p->positive_constraint = 1;
    p->compat = 1;
}

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

static uir_status upload(void **dst, const void *src, size_t bytes)
{
    TRY(uir_dmalloc(dst, bytes));
    return uir_check_hip(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
}

static uir_status asd_create(uir_batch *b, struct uir_asd **out)
{
    struct uir_asd *a = (struct uir_asd *)calloc(1, sizeof *a);
    if (!a)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
    const uint32_t S = b->n_slices;
    uir_status s = UIR_OK;
    uirk_asd_slice *sl = (uirk_asd_slice *)calloc(S, sizeof *sl);
    uirk_slice *fp = (uirk_slice *)calloc(S, sizeof *fp), *dd = (uirk_slice *)calloc(S, sizeof *dd);
    a->h_bp_sl = (uirk_slice *)calloc(S, sizeof *a->h_bp_sl);
    size_t n_items = 0, total = 0, max_nv = 1, max_tv = 1;
    for (uint32_t i = 0; i < S; ++i) {
        const uir_slice_spec *sp = &b->specs[i];
        const size_t nv = (size_t)sp->n * sp->n;
        n_items += (sp->n_dets + 127) / 128;
        total += align_up(nv * b->esi) + align_up(nv * 4);
        const size_t tiles = (size_t)((sp->n + 255) / 256) * ((sp->n + UIRK_ASD_TV_ROWS - 1) / UIRK_ASD_TV_ROWS);
        max_nv = nv > max_nv ? nv : max_nv;
        max_tv = tiles > max_tv ? tiles : max_tv;
        a->n_max = sp->n > a->n_max ? sp->n : a->n_max;
        a->a_max = sp->n_angles > a->a_max ? sp->n_angles : a->a_max;
    }
    uirk_fp_item *items = (uirk_fp_item *)calloc(n_items ? n_items : 1, sizeof *items);
    double *cost = (double *)calloc(n_items ? n_items : 1, sizeof *cost);
    if (!sl || !fp || !dd || !a->h_bp_sl || !items || !cost) {
        s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
        goto done;
    }
    {
        size_t free_b = 0;
        if ((s = uir_mem_get_info(&free_b, NULL)) != UIR_OK)
            goto done;
        if (total > free_b) {
            s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "ASD-POCS needs %zu more device bytes, %zu free", total, free_b);
            goto done;
        }
    }
    if ((s = uir_dmalloc(&a->mem, total)) != UIR_OK)
        goto done;
    a->nblk = (int)((max_nv + UIRK_ASD_CHUNK - 1) / UIRK_ASD_CHUNK);
    if ((size_t)a->nblk < max_tv)
        a->nblk = (int)max_tv;
    {
        char *p = (char *)a->mem;
        size_t k = 0;
        for (uint32_t i = 0; i < S; ++i) {
            const uir_slice_spec *sp = &b->specs[i];
            const geom_t *G = &b->geoms[b->geom_of[i]];
            const size_t nv = (size_t)sp->n * sp->n;
            sl[i].x = b->x[i];
            sl[i].x_prev = p;
            p += align_up(nv * b->esi);
            sl[i].g = (float *)p;
            p += align_up(nv * 4);
            sl[i].r = b->r[i];
            sl[i].n = sp->n;
            sl[i].n_sino = sp->n_angles * sp->n_dets;
            /* SART FP: r[k] = (b[k] - A_k x) R[k]; dd FP: r = b - A x */
            memset(&fp[i], 0, sizeof fp[i]);
            fp[i].x = b->x[i];
            fp[i].b = b->b[i];
            fp[i].r = b->r[i];
            fp[i].ray_w = G->d_ray_w;
            fp[i].trig = G->d_trig;
            fp[i].n = (int32_t)G->n;
            fp[i].n_dets = (int32_t)G->D;
            fp[i].n_angles = (int32_t)G->A;
            fp[i].axis = G->axis;
            fp[i].corr = -1.0f;
            dd[i] = fp[i];
            dd[i].ray_w = NULL;
            /* SART BP: x += relax A_k^T r[k] (pixel weight 1, no clamp) */
            uirk_slice *k2 = &a->h_bp_sl[i];
            *k2 = fp[i];
            k2->x_out = b->x[i];
            k2->b = b->r[i];
            k2->r = NULL;
            k2->ray_w = NULL;
            k2->pix_scale = 1.0f;
            k2->posc = 0;
            for (uint32_t d0 = 0; d0 < sp->n_dets; d0 += 128, ++k) {
                items[k].slice = (int32_t)i;
                items[k].d0 = (int32_t)d0;
                cost[k] = (double)sp->n;
            }
        }
    }
    a->n_items = (int)n_items;
    a->grid = uirk_plan_grid(a->n_items, 4);
    uirk_plan_xcd(cost, a->n_items, uirk_plan_num_xcd(), &a->part);
    if ((s = upload((void **)&a->d_sl, sl, S * sizeof *sl)) != UIR_OK ||
        (s = upload((void **)&a->d_fp_sl, fp, S * sizeof *fp)) != UIR_OK ||
        (s = upload((void **)&a->d_dd_sl, dd, S * sizeof *dd)) != UIR_OK ||
        (s = upload((void **)&a->d_bp_sl, a->h_bp_sl, S * sizeof *a->h_bp_sl)) != UIR_OK ||
        (s = upload((void **)&a->d_items, items, n_items * sizeof *items)) != UIR_OK ||
        (s = uir_dmalloc((void **)&a->d_state, S * sizeof(uirk_asd_state))) != UIR_OK ||
        (s = uir_dmalloc((void **)&a->d_partials, (size_t)S * a->nblk * sizeof(double))) != UIR_OK)
        goto done;
done:
    free(sl);
    free(fp);
    free(dd);
    free(items);
    free(cost);
    if (s != UIR_OK) {
        uir_asd_free(a);
        return s;
    }
    *out = a;
    return UIR_OK;
}

static int same_params(const uir_asdpocs_params *x, const uir_asdpocs_params *y)
{
    return x->beta_red == y->beta_red && x->alpha == y->alpha && x->alpha_red == y->alpha_red &&
           x->r_max == y->r_max && x->ng == y->ng && x->df_iterations == y->df_iterations &&
           x->positive_constraint == y->positive_constraint && !!x->compat == !!y->compat;
}

/* One ASD-POCS iteration of every slice (graph-capturable: kernels only). */
static uir_status launch_asd_iteration(uir_batch *b, const uir_asdpocs_params *p, uint32_t diag_iters)
{
    struct uir_asd *a = b->asd;
    hipStream_t st = b->res->stream;
    const int S = (int)b->n_slices, tb = !p->compat, nblk = a->nblk;
    const uirk_asd_params kp = {p->beta_red, p->alpha, p->alpha_red, p->r_max, tb};
#define FINISH(what)                                                                                              \
    HIP(uirk_asd_finish(a->d_state, a->d_partials, S, nblk, (what), &kp, a->d_bp_sl, a->d_diag, (int)diag_iters, \
                        st), "ASD-POCS reduction")
    HIP(uirk_asd_image(UIRK_ASD_BEGIN, b->idt, a->d_sl, S, nblk, a->d_state, tb ? 1 : 0, tb, a->d_partials, st),
        "ASD-POCS start");
    for (uint32_t pass = 0; pass < p->df_iterations; ++pass)
        for (uint32_t k = 0; k < a->a_max; ++k) {
            HIP(uirk_fp_angle(a->d_fp_sl, a->d_items, a->n_items, &a->part, b->sdt, b->idt, (int)k, a->grid, b->d_err,
                              st), "SART FP launch");
            HIP(uirk_bp_angle(a->d_bp_sl, S, (int)a->n_max, b->sdt, b->idt, (int)k, st), "SART BP launch");
        }
    HIP(uirk_asd_image(UIRK_ASD_POSC_DP, b->idt, a->d_sl, S, nblk, a->d_state, p->positive_constraint ? 1 : 0, tb,
                       a->d_partials, st), "ASD-POCS POSC");
    FINISH(UIRK_ASD_DP);
    for (int g = 0; g < UIR_FP_SHAPES; ++g) {
        const kgroup *kg = &b->fpg[g];
        if (kg->n_items)
            HIP(uirk_fp_table(a->d_dd_sl, NULL, (const uirk_fp_item *)kg->d_items, kg->n_items, &kg->part, 1, b->sdt,
                              b->idt, &kg->tile, kg->grid, b->d_err, st), "residual FP launch");
    }
    HIP(uirk_asd_sino_norm(b->sdt, a->d_sl, S, nblk, tb, a->d_partials, st), "residual norm");
    FINISH(UIRK_ASD_DD);
    const uint32_t ng = tb ? p->ng : 20; /* tvstd :573 takes 20 steps; ng is unused there */
    for (uint32_t k = 0; k < ng; ++k) {
        HIP(uirk_asd_image(UIRK_ASD_TVGRAD, b->idt, a->d_sl, S, nblk, a->d_state, 0, tb, a->d_partials, st),
            "TV gradient");
        FINISH(UIRK_ASD_G);
        HIP(uirk_asd_image(UIRK_ASD_TVSTEP, b->idt, a->d_sl, S, nblk, a->d_state, 0, tb, a->d_partials, st),
            "TV step");
    }
    HIP(uirk_asd_image(UIRK_ASD_DIFF, b->idt, a->d_sl, S, nblk, a->d_state, 0, tb, a->d_partials, st), "TV distance");
    FINISH(UIRK_ASD_DG);
#undef FINISH
    return UIR_OK;
}

uir_status uir_batch_asdpocs(uir_batch *b, uint32_t iterations, const uir_asdpocs_params *p, int use_graph)
{
    if (!b || !p)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!(p->beta_red >= 0.0f) || !(p->alpha >= 0.0f) || !(p->alpha_red >= 0.0f) || !(p->r_max >= 0.0f) ||
        !isfinite(p->beta) || p->ng > 1000 || p->df_iterations == 0)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "bad ASD-POCS parameters");
    TRY(uir_check_hip(hipSetDevice(b->res->device), "hipSetDevice"));
    if (!b->asd)
        TRY(asd_create(b, &b->asd));
    struct uir_asd *a = b->asd;
    hipStream_t st = b->res->stream;
    const uint32_t S = b->n_slices;

    /* diagnostics rows of this call (the buffer only grows, so a captured
     * graph stays valid for any shorter run) */
    if (iterations > a->diag_cap || !a->d_diag) {
        if (a->exec) { /* the captured graph holds the old buffer */
            (void)hipGraphExecDestroy(a->exec);
            (void)hipGraphDestroy(a->graph);
            a->exec = NULL;
            a->graph = NULL;
        }
        uir_dfree(a->d_diag);
        a->d_diag = NULL;
        a->diag_cap = 0;
        const uint32_t cap = iterations > 64 ? iterations : 64;
        TRY(uir_dmalloc((void **)&a->d_diag, (size_t)cap * S * 4 * sizeof(float)));
        a->diag_cap = cap;
    }
    a->diag_rows = iterations;
    HIP(hipMemsetAsync(a->d_diag, 0, (size_t)a->diag_cap * S * 4 * sizeof(float), st), "hipMemsetAsync");

    /* state: beta, the df-minimizer's relaxation (the slice's relaxation until
     * beta is a valid [0, 1] value, :459), dtgv = dd = dp = dg = 1 (:439) */
    uirk_asd_state *h = (uirk_asd_state *)calloc(S, sizeof *h);
    if (!h)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
    for (uint32_t i = 0; i < S; ++i) {
        h[i].beta = p->beta;
        h[i].relax = (p->beta >= 0.0f && p->beta <= 1.0f) ? p->beta : b->specs[i].relaxation;
        h[i].dtgv = h[i].dd = h[i].dp = h[i].dg = 1.0f;
        a->h_bp_sl[i].relax = h[i].relax;
    }
    uir_status s = uir_check_hip(hipMemcpyAsync(a->d_state, h, S * sizeof *h, hipMemcpyHostToDevice, st), "state");
    if (s == UIR_OK)
        s = uir_check_hip(hipMemcpyAsync(a->d_bp_sl, a->h_bp_sl, S * sizeof *a->h_bp_sl, hipMemcpyHostToDevice, st),
                          "descriptors");
    if (s == UIR_OK) /* the host arrays must outlive the copies */
        s = uir_check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    free(h);
    TRY(s);
    /* x = 0 and x_prev = 0 (:441-445): the first data step starts from zero in
     * both variants */
    for (uint32_t i = 0; i < S; ++i) {
        const size_t nv = (size_t)b->specs[i].n * b->specs[i].n;
        HIP(hipMemsetAsync(b->x[i], 0, nv * b->esi, st), "hipMemsetAsync");
    }
    HIP(uirk_asd_image(UIRK_ASD_BEGIN, b->idt, a->d_sl, (int)S, a->nblk, a->d_state, 1, 0, a->d_partials, st),
        "ASD-POCS start");
    if (iterations == 0)
        return UIR_OK;

    if (!use_graph) {
        for (uint32_t it = 0; it < iterations; ++it)
            TRY(launch_asd_iteration(b, p, a->diag_cap));
    } else {
        if (a->exec && !same_params(&a->graph_p, p)) {
            (void)hipGraphExecDestroy(a->exec);
            (void)hipGraphDestroy(a->graph);
            a->exec = NULL;
            a->graph = NULL;
        }
        if (!a->exec) {
            HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            uir_status s1 = launch_asd_iteration(b, p, a->diag_cap);
            hipGraph_t g = NULL;
            uir_status s2 = uir_check_hip(hipStreamEndCapture(st, &g), "hipStreamEndCapture");
            if (s1 != UIR_OK || s2 != UIR_OK) {
                if (g)
                    (void)hipGraphDestroy(g);
                return s1 != UIR_OK ? s1 : s2;
            }
            a->graph = g;
            HIP(hipGraphInstantiate(&a->exec, g, NULL, NULL, 0), "hipGraphInstantiate");
            a->graph_p = *p;
        }
        for (uint32_t it = 0; it < iterations; ++it)
            HIP(hipGraphLaunch(a->exec, st), "hipGraphLaunch");
    }
    if (p->compat) /* the reference's output is the image before the last TV steps */
        HIP(uirk_asd_image(UIRK_ASD_BEGIN, b->idt, a->d_sl, (int)S, a->nblk, a->d_state, 2, 0, a->d_partials, st),
            "ASD-POCS output");
    return UIR_OK;
}

uir_status uir_batch_asdpocs_diagnostics(uir_batch *b, uint32_t slice, float *rows4, uint32_t max_rows,
                                         uint32_t *n_rows)
{
    if (!b || slice >= b->n_slices || (!rows4 && max_rows))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "bad slice or NULL rows");
    const struct uir_asd *a = b->asd;
    const uint32_t rows = a ? a->diag_rows : 0;
    if (n_rows)
        *n_rows = rows;
    if (!rows || !max_rows)
        return UIR_OK;
    hipStream_t st = b->res->stream;
    const uint32_t S = b->n_slices, m = rows < max_rows ? rows : max_rows;
    float *all = (float *)malloc((size_t)rows * S * 4 * sizeof(float));
    if (!all)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
    uir_status s = uir_check_hip(hipMemcpyAsync(all, a->d_diag, (size_t)rows * S * 4 * sizeof(float),
                                                hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    /* rows < diag_cap: row-major [row][slice][4], the first `rows` rows */
    if (s == UIR_OK)
        s = uir_check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (s == UIR_OK)
        for (uint32_t r = 0; r < m; ++r)
            memcpy(rows4 + 4 * r, all + ((size_t)r * S + slice) * 4, 4 * sizeof(float));
    free(all);
    return s;
}
