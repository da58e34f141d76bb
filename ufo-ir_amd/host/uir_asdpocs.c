/*
 * uir_asdpocs.c -- batched ASD-POCS over the slices of a uir_batch
 * (ufo-ir-asdpocs-task.c:411-509 with a SART df-minimizer,
 * ufo-ir-sart-task.c:173-228, and tvstd :550-580).
 *
 * One iteration, for every slice at once:
 *   data step   x = 0 (reference) or x_prev = x (textbook); df_iterations SART
 *               passes: per angle k, one pass over the images that adds
 *               relax A_{k-1}^T r[k-1] and sums the FP of angle k tile by
 *               tile (uirk_sart_step), then the residual rows
 *               r[k] = (b[k] - A_k x) R[k] of every slice that has angle k
 *               (uirk_sart_rows);
 *   POSC + dp   x = max(x, 0), dp = |x - x_prev|, x_prev = x;
 *   dd          r = b - A x over all angles (the batch's FP tables), |r|;
 *   TV          ng times: g = l1_grad(x), |g|, x -= dtgv / |g| g;
 *   dg          |x - x_prev|, then the dtgv / beta updates.
 * A batch of the line model runs each SART angle as the model's FP rows and
 * BP update over every slice (uirk_sart_line: two launches, the image updated
 * in place) and its dd with the line model's FP table.
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
    void *mem;                 /* per slice: x_prev (image dtype), g (FP32), SART partial sums (FP32) */
    uirk_asd_slice *d_sl;
    uirk_sart_slice *d_sart[2]; /* fused SART step: x -> g, g -> x (g is free during the data step) */
    uirk_slice *d_dd_sl;       /* residual FP r = b - A x for dd */
    uint32_t n_max, d_max, a_max;
    int line;                  /* every slice has the line model */
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
    uir_dfree(a->d_sart[0]);
    uir_dfree(a->d_sart[1]);
    uir_dfree(a->d_dd_sl);
    uir_dfree(a->d_state);
    uir_dfree(a->d_partials);
    uir_dfree(a->d_diag);
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
    a->line = b->geoms[b->geom_of[0]].model != 0;
    uir_status s = UIR_OK;
    uirk_asd_slice *sl = (uirk_asd_slice *)calloc(S, sizeof *sl);
    uirk_sart_slice *sa = (uirk_sart_slice *)calloc(2 * (size_t)S, sizeof *sa);
    uirk_slice *dd = (uirk_slice *)calloc(S, sizeof *dd);
    size_t total = 0, max_nv = 1, max_tv = 1;
    for (uint32_t i = 0; i < S; ++i) {
        const uir_slice_spec *sp = &b->specs[i];
        const size_t nv = (size_t)sp->n * sp->n;
        const size_t part = (size_t)((sp->n + 63) / 64) * 3 * sp->n_dets;
        total += align_up(nv * b->esi) + align_up(nv * 4) + align_up(part * 4);
        const size_t tiles = (size_t)((sp->n + 255) / 256) * ((sp->n + UIRK_ASD_TV_ROWS - 1) / UIRK_ASD_TV_ROWS);
        max_nv = nv > max_nv ? nv : max_nv;
        max_tv = tiles > max_tv ? tiles : max_tv;
        a->n_max = sp->n > a->n_max ? sp->n : a->n_max;
        a->d_max = sp->n_dets > a->d_max ? sp->n_dets : a->d_max;
        a->a_max = sp->n_angles > a->a_max ? sp->n_angles : a->a_max;
    }
    if (!sl || !sa || !dd) {
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
        for (uint32_t i = 0; i < S; ++i) {
            const uir_slice_spec *sp = &b->specs[i];
            const geom_t *G = &b->geoms[b->geom_of[i]];
            const size_t nv = (size_t)sp->n * sp->n;
            const size_t part = (size_t)((sp->n + 63) / 64) * 3 * sp->n_dets;
            sl[i].x = b->x[i];
            sl[i].x_prev = p;
            p += align_up(nv * b->esi);
            sl[i].g = (float *)p;
            p += align_up(nv * 4);
            float *d_part = (float *)p;
            p += align_up(part * 4);
            sl[i].r = b->r[i];
            sl[i].n = sp->n;
            sl[i].n_sino = sp->n_angles * sp->n_dets;
            /* fused SART step: r[k] = (b[k] - A_k x) R[k], x += relax A_k^T r[k] */
            uirk_sart_slice *f = &sa[i];
            f->x_in = b->x[i];
            f->x_out = sl[i].g; /* holds image-dtype data while the data step runs */
            f->r = b->r[i];
            f->b = b->b[i];
            f->ray_w = G->d_ray_w;
            f->trig = G->d_trig;
            f->part = d_part;
            f->n = (int32_t)G->n;
            f->n_dets = (int32_t)G->D;
            f->n_angles = (int32_t)G->A;
            f->axis = G->axis;
            f->corr = -1.0f;
            if (a->line) { /* in place, with the model's angle table */
                f->x_out = b->x[i];
                f->line = G->d_line;
            }
            sa[S + i] = *f;
            sa[S + i].x_in = sl[i].g;
            sa[S + i].x_out = b->x[i];
            /* dd: r = b - A x over all angles */
            memset(&dd[i], 0, sizeof dd[i]);
            dd[i].x = b->x[i];
            dd[i].b = b->b[i];
            dd[i].r = b->r[i];
            dd[i].trig = G->d_trig;
            dd[i].n = (int32_t)G->n;
            dd[i].n_dets = (int32_t)G->D;
            dd[i].n_angles = (int32_t)G->A;
            dd[i].axis = G->axis;
            dd[i].corr = -1.0f;
            dd[i].line = a->line ? G->d_lk : NULL;
        }
    }
    if ((s = upload((void **)&a->d_sl, sl, S * sizeof *sl)) != UIR_OK ||
        (s = upload((void **)&a->d_sart[0], sa, S * sizeof *sa)) != UIR_OK ||
        (s = upload((void **)&a->d_sart[1], sa + S, S * sizeof *sa)) != UIR_OK ||
        (s = upload((void **)&a->d_dd_sl, dd, S * sizeof *dd)) != UIR_OK ||
        (s = uir_dmalloc((void **)&a->d_state, S * sizeof(uirk_asd_state))) != UIR_OK ||
        (s = uir_dmalloc((void **)&a->d_partials, (size_t)S * a->nblk * sizeof(double))) != UIR_OK)
        goto done;
done:
    free(sl);
    free(sa);
    free(dd);
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
    HIP(uirk_asd_finish(a->d_state, a->d_partials, S, nblk, (what), &kp, a->d_diag, (int)diag_iters, st),     \
        "ASD-POCS reduction")
    HIP(uirk_asd_image(UIRK_ASD_BEGIN, b->idt, a->d_sl, S, nblk, a->d_state, tb ? 1 : 0, tb, a->d_partials, st),
        "ASD-POCS start");
    /* SART passes: launch j applies the BP of angle j - 1 and sums the FP of
     * angle j, from buffer j % 2 to (j + 1) % 2 (x, g); the last launch only
     * backprojects.  An odd launch count leaves the image in g. */
    const uint32_t A = a->a_max;
    for (uint32_t pass = 0; pass < p->df_iterations && a->line; ++pass)
        for (uint32_t j = 0; j < A; ++j)
            HIP(uirk_sart_line(a->d_sart[0], a->d_state, S, (int)a->n_max, (int)a->d_max, b->sdt, b->idt, (int)j, st),
                "SART line launch");
    for (uint32_t pass = 0; pass < p->df_iterations && !a->line; ++pass) {
        for (uint32_t j = 0; j <= A; ++j) {
            const int fp = j < A ? (int)j : -1;
            HIP(uirk_sart_step(a->d_sart[j % 2], a->d_state, S, (int)a->n_max, b->sdt, b->idt, (int)j - 1, fp, b->d_err,
                               st), "SART step launch");
            if (fp >= 0)
                HIP(uirk_sart_rows(a->d_sart[j % 2], S, (int)a->d_max, b->sdt, fp, st), "SART residual launch");
        }
        if ((A + 1) % 2)
            HIP(uirk_asd_image(UIRK_ASD_BEGIN, b->idt, a->d_sl, S, nblk, a->d_state, 3, 0, a->d_partials, st),
                "SART copy");
    }
    HIP(uirk_asd_image(UIRK_ASD_POSC_DP, b->idt, a->d_sl, S, nblk, a->d_state, p->positive_constraint ? 1 : 0, tb,
                       a->d_partials, st), "ASD-POCS POSC");
    FINISH(UIRK_ASD_DP);
    for (int g = 0; g < UIR_FP_SHAPES; ++g) {
        const kgroup *kg = &b->fpg[g];
        if (kg->n_items)
            HIP(uirk_fp_table(a->d_dd_sl, NULL, (const uirk_fp_item *)kg->d_items, kg->n_items, &kg->part, kg->mode,
                              b->sdt, b->idt, &kg->tile, kg->grid, b->d_err, st), "residual FP launch");
    }
    if (b->fpl.n_items) /* the line model's table (mode 2) */
        HIP(uirk_fp_table(a->d_dd_sl, NULL, (const uirk_fp_item *)b->fpl.d_items, b->fpl.n_items, &b->fpl.part, 2,
                          b->sdt, b->idt, &b->fpl.tile, b->fpl.grid, b->d_err, st), "residual FP launch");
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
    for (uint32_t g = 1; g < b->n_geoms; ++g)
        if (b->geoms[g].model != b->geoms[0].model) /* one data step for the whole batch */
            return uir_set_error(UIR_ERR_UNSUPPORTED, "batched ASD-POCS takes slices of one projection model");
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
    }
    uir_status s = uir_check_hip(hipMemcpyAsync(a->d_state, h, S * sizeof *h, hipMemcpyHostToDevice, st), "state");
    if (s == UIR_OK) /* the host array must outlive the copy */
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
