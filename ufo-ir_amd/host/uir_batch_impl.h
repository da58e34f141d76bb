/*
 * uir_batch_impl.h -- the batch object shared by the batched methods
 * (uir_batch.c: SIRT; uir_asdpocs.c: ASD-POCS).  Not installed.
 */
#ifndef UIR_BATCH_IMPL_H
#define UIR_BATCH_IMPL_H

#include "uir_internal.h"

typedef struct {
    uint32_t n, D, A;
    float step, axis;        /* resolved axis */
    float r_min;             /* ray weight floor */
    uirk_tile tile;          /* tile shape (autotuned per geometry, or the default) */
    float *h_sin, *h_cos;    /* line model: its unit directions */
    float *d_trig, *d_ray_w; /* carved from the batch allocation */
    uir_subset *runs;        /* line model: runs of one family (direction = joseph's) */
    uint32_t n_runs;
    int32_t model;           /* 0 joseph, 1 line */
    /* line model: per-angle constants of the table kernels (8 floats) and of
     * the single-angle kernels, and the pixel weights 1 / BP(1) (N x N);
     * device copies carved from the batch allocation */
    float *h_lk;
    uirk_line_angle *h_line;
    float *d_lk, *d_pix_w;
    uirk_line_angle *d_line;
} geom_t;

#define UIR_FP_SHAPES 6 /* FP angles 16, 8, 32, with pair staging, then with single values */
#define UIR_BP_SHAPES 4

typedef struct {
    uirk_tile tile;
    void *d_items;
    int n_items, grid;
    int mode; /* launch mode: 1, 3 (the table holds slice pairs), 2 (line) or 4 (line pairs) */
    uirk_xcd_part part;
} kgroup;


struct uir_batch {
    uir_resources *res;
    uir_dtype dtype;  /* storage mode */
    int sdt, idt;     /* sinogram and image element types */
    size_t ess, esi;  /* their sizes */
    uint32_t n_slices, n_geoms;
    uir_slice_spec *specs;
    uint32_t *geom_of;
    geom_t *geoms;
    void *mem;
    size_t mem_bytes;
    void **x, **b, **r;
    uirk_slice *d_fp_sl, *d_bp_sl;
    /* one launch per tile-shape group: FP by angles per item (2 shapes), BP by
     * staged angles x tile rows (4 shapes); each group's table is slice-major */
    kgroup fpg[UIR_FP_SHAPES], bpg[UIR_BP_SHAPES];
    /* the line model's slices: one FP and one BP group (default tile, mode 2) */
    kgroup fpl, bpl;
    int *d_err;
    double fp_samples, bp_samples;
    hipGraph_t graph;
    hipGraphExec_t exec;
    uint32_t graph_iters;
    struct uir_asd *asd; /* batched ASD-POCS state (uir_asdpocs.c), created on first use */
};

void uir_asd_free(struct uir_asd *a);

#endif
