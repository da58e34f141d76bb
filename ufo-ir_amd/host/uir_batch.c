/*
 * uir_batch.c -- batched SIRT over many slices of uneven geometry.
 *
 * Layout in HBM: one allocation holds, per slice, the image x (N*N), the
 * measured sinogram b (A*D) and the weighted residual r (A*D) in the storage
 * dtype, and, per distinct geometry, the FP32 ray weights R = 1/FP(1) (A*D)
 * and the (sin, cos) table.  The pixel weight of the fast projector pair is
 * the scalar 1/A (its backprojection of ones is exactly A).  A geometry of the
 * line model adds its per-angle constants and an FP32 image of pixel weights
 * C = 1/BP(1); its slices run in the table kernels' line mode.
 *
 * One iteration = one FP launch over every slice's FP work items
 * (r = (b - A x) * R) and one BP launch over every slice's tiles
 * (x += relaxation / A * A^T r, optionally clamped at 0), i.e. the SIRT update
 * of ufo-ir-sirt-task.c:202-215.  Iterations are captured once into a hipGraph
 * and replayed.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "uir_batch_impl.h"

static int fp_shape(const uirk_tile *t)
{
    return (t->fp_angles == 8 ? 1 : (t->fp_angles == 32 ? 2 : 0)) + 3 * (t->fp_single != 0);
}
static const int k_fp_shape_angles[3] = {UIRK_FP_ANGLES, 8, 32};
static int bp_shape(const uirk_tile *t) { return (t->bp_angles == 16) * 2 + (t->bp_rows == 32); }

/* BP tile rows of a geometry in this batch: its tuned rows (the line model's
 * 64), or 32 when the batch's table would not fill the device with them */
static int bp_rows_of(const geom_t *G, int rows32)
{
    return rows32 ? 32 : (G->model ? UIRK_BP_TILE : G->tile.bp_rows);
}

#define TRY(x) do { uir_status s_ = (x); if (s_ != UIR_OK) return s_; } while (0)

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

static int same_geom(const uir_slice_spec *s, const geom_t *g)
{
    return s->n == g->n && s->n_dets == g->D && s->n_angles == g->A && s->step == g->step &&
           uir_geom_axis(s->axis_position, s->n_dets) == g->axis && s->ray_weight_min == g->r_min &&
           s->model == g->model;
}

void uir_batch_free(uir_batch *b)
{
    if (!b)
        return;
    if (b->res)
        (void)hipStreamSynchronize(b->res->stream);
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    for (uint32_t i = 0; i < b->n_geoms; ++i) {
        free(b->geoms[i].h_sin);
        free(b->geoms[i].h_cos);
        free(b->geoms[i].runs);
        free(b->geoms[i].h_lk);
        free(b->geoms[i].h_line);
    }
    free(b->geoms);
    free(b->geom_of);
    free(b->specs);
    free(b->x);
    free(b->b);
    free(b->r);
    uir_dfree(b->mem);
    uir_dfree(b->d_fp_sl);
    uir_dfree(b->d_bp_sl);
    for (int g = 0; g < UIR_FP_SHAPES; ++g)
        uir_dfree(b->fpg[g].d_items);
    for (int g = 0; g < UIR_BP_SHAPES; ++g)
        uir_dfree(b->bpg[g].d_items);
    uir_dfree(b->fpl.d_items);
    uir_dfree(b->bpl.d_items);
    uir_dfree(b->d_err);
    uir_asd_free(b->asd);
    uir_resources_free(b->res);
    free(b);
}

static uir_status upload(void **dst, const void *src, size_t bytes)
{
    TRY(uir_dmalloc(dst, bytes));
    return uir_check_hip(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
}

/* R = 1 / FP(1) for one geometry, FP(1) in FP64 (uirk_ray_weights64), into
 * d_ray_w (A x D FP32); d_trig holds the geometry's (sin, cos) table.
 * Synchronises the stream. */
uir_status uir_ray_weights(void *stream, uint32_t n, uint32_t D, uint32_t A, float axis, const float *d_trig,
                           float r_min, float *d_ray_w)
{
    TRY(uir_check_hip(uirk_ray_weights64(d_trig, (int)n, (int)D, (int)A, axis, r_min, d_ray_w, stream),
                      "ray weights"));
    return uir_check_hip(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

/* The line model's weights of one geometry from its single-angle kernels:
 * R = 1/FP(1) (0 where FP(1) = 0 or below r_min) into d_ray_w (A x D) and
 * C = 1/BP(1) (0 where BP(1) < UIRK_LINE_PIX_MIN) into d_pix_w (n x n), with
 * the geometry's angle table d_line already on the device and runs its
 * family runs.  Synchronises the stream. */
uir_status uir_line_weights(void *stream, uint32_t n, uint32_t D, uint32_t A, float r_min, const uir_subset *runs,
                            uint32_t n_runs, const uirk_line_angle *d_line, float *d_ray_w, float *d_pix_w)
{
    hipStream_t st = (hipStream_t)stream;
    const size_t nv = (size_t)n * n, ns = (size_t)A * D;
    float *ones_v = NULL, *ones_s = NULL;
    uir_status s = uir_dmalloc((void **)&ones_v, nv * 4);
    if (s == UIR_OK)
        s = uir_dmalloc((void **)&ones_s, ns * 4);
    if (s == UIR_OK)
        s = uir_check_hip(uirk_set(ones_v, nv, 1.0f, st), "set");
    if (s == UIR_OK)
        s = uir_check_hip(uirk_set(ones_s, ns, 1.0f, st), "set");
    if (s == UIR_OK)
        s = uir_check_hip(uirk_set(d_ray_w, ns, 0.0f, st), "set");
    for (uint32_t r = 0; r < n_runs && s == UIR_OK; ++r)
        s = uir_check_hip(uirk_fp_line(ones_v, d_ray_w, d_ray_w, NULL, d_line, !runs[r].direction, (int)n, (int)D,
                                       (int)runs[r].offset, (int)runs[r].n, 1.0f, 0, st),
                          "FP launch");
    if (s == UIR_OK)
        s = uir_check_hip(uirk_inv_min(d_ray_w, d_ray_w, ns, r_min, st), "inv");
    if (s == UIR_OK)
        s = uir_check_hip(uirk_set(d_pix_w, nv, 0.0f, st), "set");
    if (s == UIR_OK)
        s = uir_check_hip(uirk_bp_line(d_pix_w, d_pix_w, ones_s, NULL, 0, d_line, (int)n, (int)D, 0, (int)A, 1.0f, st),
                          "BP launch");
    if (s == UIR_OK)
        s = uir_check_hip(uirk_inv_min(d_pix_w, d_pix_w, nv, UIRK_LINE_PIX_MIN, st), "inv");
    if (s == UIR_OK)
        s = uir_check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    uir_dfree(ones_v);
    uir_dfree(ones_s);
    return s;
}

/* The line model's host tables of one geometry: its angle table h_line and
 * fast-kernel constants h_lk (8 floats per angle); the unit directions replace
 * the (sin, cos) LUT in place, and runs / n_runs become the runs of one
 * direction family each. */
uir_status uir_line_geom(uint32_t A, uint32_t n, float axis, float *sin_lut, float *cos_lut, uirk_line_angle *h_line,
                         float *h_lk, uir_subset *runs, uint32_t *n_runs)
{
    float *tu = (float *)malloc(sizeof(float) * 2 * A);
    if (!tu)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
    uir_line_table(sin_lut, cos_lut, A, n, axis, h_line, tu, h_lk);
    for (uint32_t a = 0; a < A; ++a) {
        sin_lut[a] = tu[2 * a];
        cos_lut[a] = tu[2 * a + 1];
    }
    free(tu);
    uint32_t k = 0;
    for (uint32_t a = 0; a < A; ++a) {
        const int dir = !h_line[a].family;
        if (k && runs[k - 1].direction == dir) {
            runs[k - 1].n++;
        } else {
            runs[k].offset = a;
            runs[k].n = 1;
            runs[k].direction = dir;
            ++k;
        }
    }
    *n_runs = k;
    return UIR_OK;
}

uir_status uir_batch_new(uir_resources *res, uir_dtype storage, const uir_slice_spec *specs, uint32_t n_slices,
                         uir_batch **out)
{
    if (!res || !specs || !out || n_slices == 0)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument or empty batch");
    int sdt = 0, idt = 0;
    if (!uir_storage_split(storage, &sdt, &idt))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "unknown storage dtype");
    *out = NULL;
    for (uint32_t i = 0; i < n_slices; ++i) {
        const uir_slice_spec *s = &specs[i];
        if (s->n == 0 || s->n_dets == 0 || s->n_angles == 0 || s->n > 16384 || s->n_dets > 32768 ||
            s->n_angles > 65536)
            return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "slice %u: bad geometry %ux%u, %u angles", i, s->n,
                                 s->n_dets, s->n_angles);
        if ((uint64_t)s->n * s->n * 4 >= (1ull << 32) || (uint64_t)s->n_angles * s->n_dets * 4 >= (1ull << 32))
            return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "slice %u: image or sinogram of 4 GiB or more", i);
        if (!(s->relaxation >= 0.0f && s->relaxation <= 1.0f))
            return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "slice %u: relaxation outside [0, 1]", i);
        if (!(s->step >= 0.0f) || !(s->axis_position >= -1.0f))
            return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "slice %u: bad step or axis", i);
        if (!uir_geom_active_bins(s->n_dets, uir_geom_axis(s->axis_position, s->n_dets), NULL, NULL))
            return uir_set_error(UIR_ERR_INVALID_ARGUMENT,
                                 "slice %u: axis %g leaves no active detector (needs 0.5 <= axis <= n_dets - 0.5)", i,
                                 (double)s->axis_position);
        if (!(s->ray_weight_min >= 0.0f) || !isfinite(s->ray_weight_min))
            return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "slice %u: bad ray_weight_min", i);
        if (s->model != 0 && s->model != 1)
            return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "slice %u: unknown model %d", i, s->model);
    }
    TRY(uir_check_hip(hipSetDevice(res->device), "hipSetDevice"));
    uir_batch *b = (uir_batch *)calloc(1, sizeof *b);
    if (!b)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
    b->res = uir_resources_retain(res);
    b->dtype = storage;
    b->sdt = sdt;
    b->idt = idt;
    b->ess = uir_dtype_size(sdt);
    b->esi = uir_dtype_size(idt);
    b->n_slices = n_slices;
    b->specs = (uir_slice_spec *)malloc(sizeof *specs * n_slices);
    b->geom_of = (uint32_t *)calloc(n_slices, sizeof(uint32_t));
    b->geoms = (geom_t *)calloc(n_slices, sizeof(geom_t));
    b->x = (void **)calloc(n_slices, sizeof(void *));
    b->b = (void **)calloc(n_slices, sizeof(void *));
    b->r = (void **)calloc(n_slices, sizeof(void *));
    uir_status s = UIR_OK;
    uirk_slice *fp_sl = NULL, *bp_sl = NULL;
    char *lead = NULL; /* slice i leads a pair with slice i + 1 */
    uirk_fp_vec fv[UIR_FP_SHAPES] = {{0}}, fvl = {0};
    uirk_bp_vec bv[UIR_BP_SHAPES] = {{0}}, bvl = {0};
    if (!b->specs || !b->geom_of || !b->geoms || !b->x || !b->b || !b->r) {
        s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
        goto fail;
    }
    memcpy(b->specs, specs, sizeof *specs * n_slices);

    /* distinct geometries: LUTs and direction runs as the projector builds them */
    for (uint32_t i = 0; i < n_slices; ++i) {
        uint32_t g = 0;
        while (g < b->n_geoms && !same_geom(&specs[i], &b->geoms[g]))
            ++g;
        if (g == b->n_geoms) {
            geom_t *G = &b->geoms[b->n_geoms++];
            G->n = specs[i].n;
            G->D = specs[i].n_dets;
            G->A = specs[i].n_angles;
            G->step = specs[i].step;
            G->axis = uir_geom_axis(specs[i].axis_position, specs[i].n_dets);
            G->r_min = specs[i].ray_weight_min;
            G->model = specs[i].model;
            G->tile = G->model ? (uirk_tile)UIRK_TILE_DEFAULT : uir_tile_for(&specs[i], storage);
            G->h_sin = (float *)malloc(sizeof(float) * G->A);
            G->h_cos = (float *)malloc(sizeof(float) * G->A);
            G->runs = (uir_subset *)malloc(sizeof(uir_subset) * G->A);
            if (!G->h_sin || !G->h_cos || !G->runs) {
                s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
                goto fail;
            }
            uir_geom_lut(G->A, G->step, G->h_sin, G->h_cos);
            if (G->model) {
                G->h_lk = (float *)malloc(sizeof(float) * 8 * G->A);
                G->h_line = (uirk_line_angle *)malloc(sizeof(uirk_line_angle) * G->A);
                if (!G->h_lk || !G->h_line) {
                    s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
                    goto fail;
                }
                if ((s = uir_line_geom(G->A, G->n, G->axis, G->h_sin, G->h_cos, G->h_line, G->h_lk, G->runs,
                                       &G->n_runs)) != UIR_OK)
                    goto fail;
            } else {
                G->n_runs = uir_geom_runs(G->h_sin, G->h_cos, G->A, G->runs);
            }
        }
        b->geom_of[i] = g;
    }

    /* one allocation for everything resident */
    size_t total = 0;
    for (uint32_t i = 0; i < n_slices; ++i) {
        const size_t nv = (size_t)specs[i].n * specs[i].n, ns = (size_t)specs[i].n_angles * specs[i].n_dets;
        total += align_up(nv * b->esi) + 2 * align_up(ns * b->ess);
    }
    for (uint32_t g = 0; g < b->n_geoms; ++g) {
        const geom_t *G = &b->geoms[g];
        total += align_up((size_t)G->A * G->D * 4) + align_up((size_t)G->A * 8);
        if (G->model)
            total += align_up((size_t)G->A * 32) + align_up((size_t)G->n * G->n * 4) +
                     align_up((size_t)G->A * sizeof(uirk_line_angle));
    }
    {
        size_t free_b = 0;
        if ((s = uir_mem_get_info(&free_b, NULL)) != UIR_OK)
            goto fail;
        if (total > free_b) {
            s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "batch needs %zu device bytes, %zu free", total, free_b);
            goto fail;
        }
    }
    if ((s = uir_dmalloc(&b->mem, total)) != UIR_OK)
        goto fail;
    b->mem_bytes = total;
    {
        char *p = (char *)b->mem;
        for (uint32_t i = 0; i < n_slices; ++i) {
            const size_t nv = (size_t)specs[i].n * specs[i].n, ns = (size_t)specs[i].n_angles * specs[i].n_dets;
            b->x[i] = p;
            p += align_up(nv * b->esi);
            b->b[i] = p;
            p += align_up(ns * b->ess);
            b->r[i] = p;
            p += align_up(ns * b->ess);
        }
        for (uint32_t g = 0; g < b->n_geoms; ++g) {
            geom_t *G = &b->geoms[g];
            G->d_ray_w = (float *)p;
            p += align_up((size_t)G->A * G->D * 4);
            G->d_trig = (float *)p;
            p += align_up((size_t)G->A * 8);
            if (G->model) {
                G->d_lk = (float *)p;
                p += align_up((size_t)G->A * 32);
                G->d_pix_w = (float *)p;
                p += align_up((size_t)G->n * G->n * 4);
                G->d_line = (uirk_line_angle *)p;
                p += align_up((size_t)G->A * sizeof(uirk_line_angle));
            }
        }
    }
    if ((s = uir_check_hip(hipMemset(b->mem, 0, total), "hipMemset")) != UIR_OK)
        goto fail;
    if ((s = uir_dmalloc((void **)&b->d_err, sizeof(int))) != UIR_OK ||
        (s = uir_check_hip(hipMemset(b->d_err, 0, sizeof(int)), "hipMemset")) != UIR_OK)
        goto fail;
    for (uint32_t g = 0; g < b->n_geoms; ++g) {
        geom_t *G = &b->geoms[g];
        float *trig = (float *)malloc(sizeof(float) * 2 * G->A);
        if (!trig) {
            s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
            goto fail;
        }
        uirk_plan_trig(G->h_sin, G->h_cos, G->A, trig);
        s = uir_check_hip(hipMemcpy(G->d_trig, trig, sizeof(float) * 2 * G->A, hipMemcpyHostToDevice), "hipMemcpy");
        free(trig);
        if (s != UIR_OK)
            goto fail;
        if (G->model) {
            if ((s = uir_check_hip(hipMemcpy(G->d_lk, G->h_lk, sizeof(float) * 8 * G->A, hipMemcpyHostToDevice),
                                   "hipMemcpy")) != UIR_OK ||
                (s = uir_check_hip(hipMemcpy(G->d_line, G->h_line, sizeof(uirk_line_angle) * G->A,
                                             hipMemcpyHostToDevice), "hipMemcpy")) != UIR_OK ||
                (s = uir_line_weights(b->res->stream, G->n, G->D, G->A, G->r_min, G->runs, G->n_runs, G->d_line,
                                      G->d_ray_w, G->d_pix_w)) != UIR_OK)
                goto fail;
        } else if ((s = uir_ray_weights(b->res->stream, G->n, G->D, G->A, G->axis, G->d_trig, G->r_min,
                                        G->d_ray_w)) != UIR_OK) {
            goto fail;
        }
    }

    /* descriptors and work tables */
    fp_sl = (uirk_slice *)calloc(n_slices, sizeof *fp_sl);
    bp_sl = (uirk_slice *)calloc(n_slices, sizeof *bp_sl);
    lead = (char *)calloc(n_slices, 1);
    if (!fp_sl || !bp_sl || !lead) {
        s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
        goto fail;
    }
    /* BP tiles of 32 rows when the table at the tuned rows would not fill the
     * device; a pixel's BP result does not depend on its tile's rows */
    size_t bp_items = 0;
    for (uint32_t i = 0; i < n_slices; ++i) {
        const geom_t *G = &b->geoms[b->geom_of[i]];
        const int rows = bp_rows_of(G, 0);
        bp_items += (size_t)((G->n + UIRK_BP_TILE - 1) / UIRK_BP_TILE) * ((G->n + rows - 1) / rows);
    }
    const int rows32 = uirk_plan_bp_rows32(bp_items);
    /* consecutive slices of one geometry pair up in the BP (modes 3 and 4),
     * per model if its paired table still fills the device */
    for (int32_t model = 0; model < 2; ++model) {
        size_t paired_items = 0;
        for (uint32_t i = 0; i < n_slices; ++i) {
            const geom_t *G = &b->geoms[b->geom_of[i]];
            const int rows = bp_rows_of(G, rows32);
            const size_t tiles = (size_t)((G->n + UIRK_BP_TILE - 1) / UIRK_BP_TILE) * ((G->n + rows - 1) / rows);
            if (G->model != model)
                continue;
            paired_items += tiles;
            if (i + 1 < n_slices && b->geom_of[i + 1] == b->geom_of[i]) {
                lead[i] = 1;
                ++i;
            }
        }
        if (paired_items < (size_t)uirk_plan_pair_min_items())
            for (uint32_t i = 0; i < n_slices; ++i)
                lead[i] = b->geoms[b->geom_of[i]].model == model ? 0 : lead[i];
    }
    for (uint32_t i = 0; i < n_slices; ++i) {
        const geom_t *G = &b->geoms[b->geom_of[i]];
        uirk_slice *f = &fp_sl[i], *k = &bp_sl[i];
        f->x = b->x[i];
        f->b = b->b[i];
        f->r = b->r[i];
        f->ray_w = G->d_ray_w;
        f->trig = G->d_trig;
        f->n = (int32_t)G->n;
        f->n_dets = (int32_t)G->D;
        f->n_angles = (int32_t)G->A;
        f->axis = G->axis;
        f->corr = -1.0f;
        *k = *f;
        k->x = b->x[i];
        k->x_out = b->x[i];
        k->b = b->r[i];
        k->r = NULL;
        k->ray_w = NULL;
        k->relax = specs[i].relaxation;
        k->pix_scale = 1.0f / (float)G->A;
        k->posc = specs[i].positive_constraint ? 1 : 0;
        k->n_runs = 1;
        k->run_off[0] = 0;
        k->run_n[0] = G->A;
        if (G->model) { /* line: its constants, and per-pixel weights in place of 1/A */
            f->line = G->d_lk;
            k->line = G->d_lk;
            k->pix_w = G->d_pix_w;
            k->pix_scale = 1.0f;
            const int second = i > 0 && lead[i - 1];
            const size_t b0 = bvl.n;
            if (uirk_plan_fp_line(&fvl, (int)i, (int)G->n, (int)G->D, G->axis, G->h_sin, G->h_cos, G->runs,
                                  G->n_runs, UIRK_FP_ANGLES, UIRK_FP_WMAX) ||
                (!second && uirk_plan_bp(&bvl, (int)i, (int)G->n, bp_rows_of(G, rows32)))) {
                s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "work table");
                goto fail;
            }
            for (size_t j = b0; j < bvl.n && lead[i]; ++j)
                bvl.v[j].pair = 1;
        } else {
            /* BP slice pairs: slice i + 1 of the same geometry joins slice i's
             * BP items (pair set) */
            const int second = i > 0 && lead[i - 1];
            uirk_tile bt = G->tile;
            bt.bp_rows = bp_rows_of(G, rows32);
            uirk_bp_vec *bvg = &bv[bp_shape(&bt)];
            const size_t b0 = bvg->n;
            if (uirk_plan_fp(&fv[fp_shape(&G->tile)], (int)i, (int)G->n, (int)G->D, G->axis, G->h_sin, G->h_cos,
                             G->runs, G->n_runs, G->tile.fp_angles, UIRK_TILE_FP_WMAX(G->tile)) ||
                (!second && uirk_plan_bp(bvg, (int)i, (int)G->n, bt.bp_rows))) {
                s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "work table");
                goto fail;
            }
            for (size_t j = b0; j < bvg->n && lead[i]; ++j)
                bvg->v[j].pair = 1;
        }
        double m = 0;
        for (uint32_t a = 0; a < G->A; ++a)
            m += fabs(G->h_sin[a]) > fabs(G->h_cos[a]) ? fabs(G->h_sin[a]) : fabs(G->h_cos[a]);
        b->fp_samples += m * (double)G->n * G->n;
        b->bp_samples += (double)G->A * G->n * G->n;
    }
    /* Tables stay slice-major; each XCD takes a contiguous, cost-balanced range,
     * so its L2 holds the images and sinograms of a few slices only. */
    {
        size_t nmax = fvl.n > bvl.n ? fvl.n : bvl.n;
        nmax = nmax > 1 ? nmax : 1;
        for (int g = 0; g < UIR_FP_SHAPES; ++g)
            nmax = fv[g].n > nmax ? fv[g].n : nmax;
        for (int g = 0; g < UIR_BP_SHAPES; ++g)
            nmax = bv[g].n > nmax ? bv[g].n : nmax;
        double *cost = (double *)malloc(sizeof(double) * nmax);
        if (!cost) {
            s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
            goto fail;
        }
        for (int g = 0; g < UIR_FP_SHAPES && s == UIR_OK; ++g) {
            kgroup *k = &b->fpg[g];
            k->tile = (uirk_tile)UIRK_TILE_DEFAULT;
            k->tile.fp_angles = k_fp_shape_angles[g % 3];
            k->tile.fp_single = g / 3;
            k->mode = 1;
            for (size_t i = 0; i < fv[g].n; ++i)
                cost[i] = uirk_fp_item_cost(&fv[g].v[i]);
            uirk_plan_xcd(cost, (int)fv[g].n, uirk_plan_num_xcd(), &k->part);
            k->n_items = (int)fv[g].n;
            k->grid = uirk_plan_grid_xcd(&k->part, k->n_items);
            if (k->n_items)
                s = upload(&k->d_items, fv[g].v, fv[g].n * sizeof(uirk_fp_item));
        }
        for (int g = 0; g < UIR_BP_SHAPES && s == UIR_OK; ++g) {
            kgroup *k = &b->bpg[g];
            k->tile = (uirk_tile)UIRK_TILE_DEFAULT;
            k->tile.bp_angles = g >= 2 ? 16 : 32;
            k->tile.bp_rows = (g & 1) ? 32 : UIRK_BP_TILE;
            k->mode = 1;
            /* an item costs its angles: in a launch with pairs, an unpaired
             * item runs the pair kernel and costs what a pair does */
            for (size_t i = 0; i < bv[g].n; ++i) {
                cost[i] = (double)specs[bv[g].v[i].slice].n_angles;
                k->mode = bv[g].v[i].pair ? 3 : k->mode;
            }
            uirk_plan_xcd(cost, (int)bv[g].n, uirk_plan_num_xcd(), &k->part);
            k->n_items = (int)bv[g].n;
            k->grid = uirk_plan_grid_xcd(&k->part, k->n_items);
            if (k->n_items)
                s = upload(&k->d_items, bv[g].v, bv[g].n * sizeof(uirk_bp_item));
        }
        if (s == UIR_OK) { /* the line model's groups */
            kgroup *k = &b->fpl;
            k->tile = (uirk_tile)UIRK_TILE_DEFAULT;
            for (size_t i = 0; i < fvl.n; ++i)
                cost[i] = uirk_fp_item_cost(&fvl.v[i]);
            uirk_plan_xcd(cost, (int)fvl.n, uirk_plan_num_xcd(), &k->part);
            k->n_items = (int)fvl.n;
            k->grid = uirk_plan_grid_xcd(&k->part, k->n_items);
            if (k->n_items)
                s = upload(&k->d_items, fvl.v, fvl.n * sizeof(uirk_fp_item));
        }
        if (s == UIR_OK) {
            kgroup *k = &b->bpl;
            k->tile = (uirk_tile)UIRK_TILE_DEFAULT;
            k->tile.bp_rows = rows32 ? 32 : UIRK_BP_TILE;
            k->mode = 2;
            for (size_t i = 0; i < bvl.n; ++i) {
                cost[i] = (double)specs[bvl.v[i].slice].n_angles;
                k->mode = bvl.v[i].pair ? 4 : k->mode;
            }
            uirk_plan_xcd(cost, (int)bvl.n, uirk_plan_num_xcd(), &k->part);
            k->n_items = (int)bvl.n;
            k->grid = uirk_plan_grid_xcd(&k->part, k->n_items);
            if (k->n_items)
                s = upload(&k->d_items, bvl.v, bvl.n * sizeof(uirk_bp_item));
        }
        free(cost);
        if (s != UIR_OK)
            goto fail;
    }
    if ((s = upload((void **)&b->d_fp_sl, fp_sl, n_slices * sizeof *fp_sl)) != UIR_OK ||
        (s = upload((void **)&b->d_bp_sl, bp_sl, n_slices * sizeof *bp_sl)) != UIR_OK)
        goto fail;
    free(fp_sl);
    free(bp_sl);
    free(lead);
    for (int g = 0; g < UIR_FP_SHAPES; ++g)
        free(fv[g].v);
    for (int g = 0; g < UIR_BP_SHAPES; ++g)
        free(bv[g].v);
    free(fvl.v);
    free(bvl.v);
    *out = b;
    return UIR_OK;
fail:
    free(fp_sl);
    free(bp_sl);
    free(lead);
    for (int g = 0; g < UIR_FP_SHAPES; ++g)
        free(fv[g].v);
    for (int g = 0; g < UIR_BP_SHAPES; ++g)
        free(bv[g].v);
    free(fvl.v);
    free(bvl.v);
    uir_batch_free(b);
    return s;
}

uir_status uir_batch_set_sinogram(uir_batch *b, uint32_t slice, const float *sino, int on_device)
{
    if (!b || !sino || slice >= b->n_slices)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "bad slice or NULL sinogram");
    const size_t ns = (size_t)b->specs[slice].n_angles * b->specs[slice].n_dets;
    hipStream_t st = b->res->stream;
    if (on_device)
        return uir_check_hip(uirk_convert(sino, UIR_DTYPE_F32, b->b[slice], b->sdt, ns, st), "convert");
    float *tmp = NULL;
    TRY(uir_dmalloc((void **)&tmp, ns * 4));
    uir_status s = uir_check_hip(hipMemcpyAsync(tmp, sino, ns * 4, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    if (s == UIR_OK)
        s = uir_check_hip(uirk_convert(tmp, UIR_DTYPE_F32, b->b[slice], b->sdt, ns, st), "convert");
    if (s == UIR_OK)
        s = uir_check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    uir_dfree(tmp);
    return s;
}

uir_status uir_batch_reset(uir_batch *b)
{
    if (!b)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "batch is NULL");
    for (uint32_t i = 0; i < b->n_slices; ++i) {
        const size_t nv = (size_t)b->specs[i].n * b->specs[i].n;
        TRY(uir_check_hip(hipMemsetAsync(b->x[i], 0, nv * b->esi, b->res->stream), "hipMemsetAsync"));
    }
    return UIR_OK;
}

static uir_status launch_fp_groups(uir_batch *b)
{
    for (int g = 0; g < UIR_FP_SHAPES; ++g) {
        const kgroup *k = &b->fpg[g];
        if (k->n_items)
            TRY(uir_check_hip(uirk_fp_table(b->d_fp_sl, NULL, (const uirk_fp_item *)k->d_items, k->n_items, &k->part, k->mode,
                                            b->sdt, b->idt, &k->tile, k->grid, b->d_err, b->res->stream),
                              "FP launch"));
    }
    if (b->fpl.n_items)
        TRY(uir_check_hip(uirk_fp_table(b->d_fp_sl, NULL, (const uirk_fp_item *)b->fpl.d_items, b->fpl.n_items,
                                        &b->fpl.part, 2, b->sdt, b->idt, &b->fpl.tile, b->fpl.grid, b->d_err,
                                        b->res->stream),
                          "FP launch"));
    return UIR_OK;
}

static uir_status launch_bp_groups(uir_batch *b)
{
    for (int g = 0; g < UIR_BP_SHAPES; ++g) {
        const kgroup *k = &b->bpg[g];
        if (k->n_items)
            TRY(uir_check_hip(uirk_bp_table(b->d_bp_sl, NULL, (const uirk_bp_item *)k->d_items, k->n_items, &k->part, k->mode,
                                            b->sdt, b->idt, &k->tile, k->grid, b->d_err, b->res->stream),
                              "BP launch"));
    }
    if (b->bpl.n_items)
        TRY(uir_check_hip(uirk_bp_table(b->d_bp_sl, NULL, (const uirk_bp_item *)b->bpl.d_items, b->bpl.n_items,
                                        &b->bpl.part, b->bpl.mode, b->sdt, b->idt, &b->bpl.tile, b->bpl.grid, b->d_err,
                                        b->res->stream),
                          "BP launch"));
    return UIR_OK;
}

static uir_status launch_iteration(uir_batch *b)
{
    TRY(launch_fp_groups(b));
    return launch_bp_groups(b);
}

uir_status uir_batch_iterate(uir_batch *b, uint32_t iterations, int use_graph)
{
    if (!b)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (iterations == 0)
        return UIR_OK;
    TRY(uir_check_hip(hipSetDevice(b->res->device), "hipSetDevice"));
    if (!use_graph) {
        for (uint32_t i = 0; i < iterations; ++i)
            TRY(launch_iteration(b));
        return UIR_OK;
    }
    if (!b->exec || b->graph_iters != iterations) {
        if (b->exec) {
            (void)hipGraphExecDestroy(b->exec);
            b->exec = NULL;
        }
        if (b->graph) {
            (void)hipGraphDestroy(b->graph);
            b->graph = NULL;
        }
        hipStream_t st = b->res->stream;
        TRY(uir_check_hip(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"));
        uir_status s = UIR_OK;
        for (uint32_t i = 0; i < iterations && s == UIR_OK; ++i)
            s = launch_iteration(b);
        hipGraph_t g = NULL;
        uir_status s2 = uir_check_hip(hipStreamEndCapture(st, &g), "hipStreamEndCapture");
        if (s != UIR_OK || s2 != UIR_OK) {
            if (g)
                (void)hipGraphDestroy(g);
            return s != UIR_OK ? s : s2;
        }
        b->graph = g;
        TRY(uir_check_hip(hipGraphInstantiate(&b->exec, g, NULL, NULL, 0), "hipGraphInstantiate"));
        b->graph_iters = iterations;
    }
    return uir_check_hip(hipGraphLaunch(b->exec, b->res->stream), "hipGraphLaunch");
}

uir_status uir_batch_get_image(uir_batch *b, uint32_t slice, float *dst, int on_device)
{
    if (!b || !dst || slice >= b->n_slices)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "bad slice or NULL destination");
    const size_t nv = (size_t)b->specs[slice].n * b->specs[slice].n;
    hipStream_t st = b->res->stream;
    int err = 0;
    TRY(uir_check_hip(hipMemcpyAsync(&err, b->d_err, sizeof err, hipMemcpyDeviceToHost, st), "flag read"));
    TRY(uir_check_hip(hipStreamSynchronize(st), "hipStreamSynchronize"));
    if (err)
        return uir_set_error(UIR_ERR_DEVICE, "projector work exceeded a staging limit (flag %d: 1 FP window, 2 BP "
                             "window, 4 FP item angles, 8 SART candidate rays)", err);
    if (on_device) {
        TRY(uir_check_hip(uirk_convert(b->x[slice], b->idt, dst, UIR_DTYPE_F32, nv, st), "convert"));
        return uir_check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    float *tmp = NULL;
    TRY(uir_dmalloc((void **)&tmp, nv * 4));
    uir_status s = uir_check_hip(uirk_convert(b->x[slice], b->idt, tmp, UIR_DTYPE_F32, nv, st), "convert");
    if (s == UIR_OK)
        s = uir_check_hip(hipMemcpyAsync(dst, tmp, nv * 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (s == UIR_OK)
        s = uir_check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    uir_dfree(tmp);
    return s;
}

uir_status uir_batch_get_info(const uir_batch *b, uir_batch_info *info)
{
    if (!b || !info)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    memset(info, 0, sizeof *info);
    uint64_t nf = 0, nb = 0;
    for (int g = 0; g < UIR_FP_SHAPES; ++g) {
        nf += (uint64_t)b->fpg[g].n_items;
        info->fp_grid += (uint32_t)b->fpg[g].grid * (b->fpg[g].n_items > 0);
    }
    for (int g = 0; g < UIR_BP_SHAPES; ++g) {
        nb += (uint64_t)b->bpg[g].n_items;
        info->bp_grid += (uint32_t)b->bpg[g].grid * (b->bpg[g].n_items > 0);
    }
    nf += (uint64_t)b->fpl.n_items;
    nb += (uint64_t)b->bpl.n_items;
    info->fp_grid += (uint32_t)b->fpl.grid * (b->fpl.n_items > 0);
    info->bp_grid += (uint32_t)b->bpl.grid * (b->bpl.n_items > 0);
    info->device_bytes = b->mem_bytes + nf * sizeof(uirk_fp_item) + nb * sizeof(uirk_bp_item) +
                         2ull * b->n_slices * sizeof(uirk_slice);
    info->n_slices = b->n_slices;
    info->n_geometries = b->n_geoms;
    info->fp_items = (uint32_t)nf;
    info->bp_items = (uint32_t)nb;
    info->fp_samples = b->fp_samples;
    info->bp_samples = b->bp_samples;
    return UIR_OK;
}

uir_status uir_batch_time_kernels(uir_batch *b, uint32_t reps, float *fp_ms, float *bp_ms)
{
    if (!b || reps == 0)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "bad arguments");
    hipStream_t st = b->res->stream;
    hipEvent_t e0, e1, e2;
    TRY(uir_check_hip(hipEventCreate(&e0), "hipEventCreate"));
    TRY(uir_check_hip(hipEventCreate(&e1), "hipEventCreate"));
    TRY(uir_check_hip(hipEventCreate(&e2), "hipEventCreate"));
    uir_status s = uir_check_hip(hipEventRecord(e0, st), "hipEventRecord");
    for (uint32_t i = 0; i < reps && s == UIR_OK; ++i)
        s = launch_fp_groups(b);
    if (s == UIR_OK)
        s = uir_check_hip(hipEventRecord(e1, st), "hipEventRecord");
    for (uint32_t i = 0; i < reps && s == UIR_OK; ++i)
        s = launch_bp_groups(b);
    if (s == UIR_OK)
        s = uir_check_hip(hipEventRecord(e2, st), "hipEventRecord");
    if (s == UIR_OK)
        s = uir_check_hip(hipEventSynchronize(e2), "hipEventSynchronize");
    float a = 0, c = 0;
    if (s == UIR_OK)
        s = uir_check_hip(hipEventElapsedTime(&a, e0, e1), "hipEventElapsedTime");
    if (s == UIR_OK)
        s = uir_check_hip(hipEventElapsedTime(&c, e1, e2), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (s == UIR_OK) {
        if (fp_ms) *fp_ms = a / reps;
        if (bp_ms) *bp_ms = c / reps;
    }
    return s;
}
