/*
 * uir_tasks.c -- the task objects of libuir: the parallel-beam projector and the
 * SIRT, SART, ASD-POCS and split-Bregman TV methods, with the reference's
 * property surface, requisition and process semantics (src/tasks/ufo-ir-*-task.c and
 * src/core/ufo-ir-{state-dependent,projector,method}-task.c).
 *
 * Each process() enqueues on the task's resources stream; reductions the
 * reference ran on the host (L1 norms, the CGS "dot") run on the device and
 * synchronise the stream to hand their scalar to the host loop, as the
 * reference's ufo_buffer_get_host_array did.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "uir_internal.h"

typedef enum { K_PROJECTOR = 0, K_SIRT, K_SART, K_ASDPOCS, K_SBTV } task_kind;
typedef enum { P_FLOAT = UIR_PROP_FLOAT, P_UINT = UIR_PROP_UINT, P_BOOL = UIR_PROP_BOOL, P_STRING = UIR_PROP_STRING,
               P_OBJECT = UIR_PROP_OBJECT } prop_type;

#define N_TMP 32

struct uir_task {
    int refcount;
    task_kind kind;
    uir_resources *res;

    /* projector (ufo-ir-projector-task.c:68-110, state-dependent :75-80,
     * parallel-projector :123-137) */
    float axis_position, step, relaxation, correction_scale;
    int is_forward;
    char model[64];
    uint32_t angles_num, image_size, detectors_num;
    int first_run;
    uint32_t n_dets;
    float *h_sin, *h_cos, *d_sin, *d_cos;
    uint32_t lut_n;
    uir_subset *runs;
    uint32_t n_runs;
    /* work tables of the table-driven kernels for the current geometry */
    uint32_t plan_n, plan_dets, plan_angles;
    float plan_axis;
    uirk_fp_item *d_fp_items;
    uirk_bp_item *d_bp_items;
    int n_fp_items, n_bp_items;
    uirk_xcd_part fp_part, bp_part;
    float *d_trig;
    int *d_err;
    /* the line model (model "line"): per-angle constants of the current
     * geometry (host copy and device table) and the angles' family runs */
    int line;
    uint32_t lut_gen, line_gen, line_n, line_angles;
    float line_axis;
    uirk_line_angle *h_line, *d_line;
    uir_subset *line_runs;
    uint32_t n_line_runs;
    /* the line model on the table-driven fast path (mode 2): unit-direction
     * trig pairs, per-angle constants, work tables and their partitions */
    float *d_ltrig, *d_lk;
    uirk_fp_item *d_lfp_items;
    uirk_bp_item *d_lbp_items;
    int n_lfp_items, n_lbp_items;
    uirk_xcd_part lfp_part, lbp_part;

    /* methods (ufo-ir-method-task.c:74-85 and the method tasks) */
    uir_task *projector;
    uint32_t num_iterations;
    float relaxation_factor;
    int positive_constraint;
    float beta, beta_red, alpha, alpha_red, r_max;
    uint32_t ng;
    uir_task *df_minimizer;
    float *diag;
    uint32_t diag_rows;
    float lambda, mu;
    int compat; /* asdpocs / sbtv: 1 = the reference's arithmetic, 0 = the textbook variant */

    /* cached device temporaries (ufo_buffer_dup in the reference) */
    void *tmp[N_TMP];
    size_t tmp_bytes[N_TMP];

    /* one SART sweep captured as a hipGraph, valid for sart_key */
    hipGraph_t sart_graph;
    hipGraphExec_t sart_exec;
    struct {
        const void *x, *b, *tmp, *ray_w, *sin, *cos, *starts, *line;
        uint32_t n, D, A;
        float axis, relax, corr;
    } sart_key;
};

typedef struct {
    const char *name; /* canonical: '-' separated */
    prop_type type;
    size_t offset;
    double min, max;
} prop_def;

#define OFF(f) offsetof(struct uir_task, f)

static const prop_def g_projector_props[] = {
    {"axis-position", P_FLOAT, OFF(axis_position), -1.0, FLT_MAX},   /* ufo-ir-projector-task.c:68-73 */
    {"step", P_FLOAT, OFF(step), 0.0, FLT_MAX},                      /* :75-80 */
    {"relaxation", P_FLOAT, OFF(relaxation), FLT_MIN, FLT_MAX},      /* :82-87 */
    {"correction-scale", P_FLOAT, OFF(correction_scale), FLT_MIN, FLT_MAX}, /* :88-93 */
    {"is-forward", P_BOOL, OFF(is_forward), 0, 1},                   /* ufo-ir-state-dependent-task.c:75-80 */
    {"model", P_STRING, OFF(model), 0, 0},                           /* ufo-ir-parallel-projector-task.c:123-128 */
    {"angles-num", P_UINT, OFF(angles_num), 0, 4294967295.0},        /* :132-137 */
    {"image-size", P_UINT, OFF(image_size), 0, 1048576},             /* extension: N x N image edge (0 = n_dets) */
    {"detectors-num", P_UINT, OFF(detectors_num), 0, 1048576},       /* extension: FP output width (0 = input width) */
    {NULL, 0, 0, 0, 0}};

#define METHOD_PROPS                                                               \
    {"num-iterations", P_UINT, OFF(num_iterations), 1, 4294967295.0}, /* ufo-ir-method-task.c:74-79 */ \
    {"projector", P_OBJECT, OFF(projector), 0, 0}                      /* :80-85 */

static const prop_def g_sirt_props[] = {
    METHOD_PROPS,
    {"relaxation-factor", P_FLOAT, OFF(relaxation_factor), 0.0, 1.0}, /* ufo-ir-sirt-task.c:72-77 */
    {"positive-constraint", P_BOOL, OFF(positive_constraint), 0, 1},   /* extension: clamp x >= 0 per update */
    {NULL, 0, 0, 0, 0}};

static const prop_def g_sart_props[] = {
    METHOD_PROPS,
    {"relaxation-factor", P_FLOAT, OFF(relaxation_factor), 0.0, 1.0}, /* ufo-ir-sart-task.c:76-81 */
    {NULL, 0, 0, 0, 0}};

static const prop_def g_asdpocs_props[] = {
    METHOD_PROPS,
    {"beta", P_FLOAT, OFF(beta), 0.0, FLT_MAX},                       /* ufo-ir-asdpocs-task.c:104-109 */
    {"beta-red", P_FLOAT, OFF(beta_red), 0.0, FLT_MAX},               /* :111-116 */
    {"ng", P_UINT, OFF(ng), 0, 1000},                                 /* :118-123 */
    {"alpha", P_FLOAT, OFF(alpha), 0.0, FLT_MAX},                     /* :125-130 */
    {"alpha-red", P_FLOAT, OFF(alpha_red), 0.0, FLT_MAX},             /* :132-137 */
    {"r-max", P_FLOAT, OFF(r_max), 0.0, FLT_MAX},                     /* :139-144 */
    {"positive-constraint", P_BOOL, OFF(positive_constraint), 0, 1},  /* :146-151 */
    {"df-minimizer", P_OBJECT, OFF(df_minimizer), 0, 0},              /* :153-158 */
    {"compat", P_BOOL, OFF(compat), 0, 1},                            /* extension: 0 = ASD-POCS as published */
    {NULL, 0, 0, 0, 0}};

static const prop_def g_sbtv_props[] = {
    METHOD_PROPS,
    {"lambda", P_FLOAT, OFF(lambda), 0.0, FLT_MAX},                   /* ufo-ir-sbtv-task.c:90-95 */
    {"mu", P_FLOAT, OFF(mu), 0.0, FLT_MAX},                           /* :96-101 */
    {"compat", P_BOOL, OFF(compat), 0, 1},                            /* extension: 0 = CGS with a true inner product */
    {NULL, 0, 0, 0, 0}};

static const char *g_plugin_names[] = {"parallel-projector", "sirt", "sart", "asdpocs", "sbtv"};

static const prop_def *props_of(task_kind k)
{
    switch (k) {
    case K_PROJECTOR: return g_projector_props;
    case K_SIRT: return g_sirt_props;
    case K_SART: return g_sart_props;
    case K_ASDPOCS: return g_asdpocs_props;
    default: return g_sbtv_props;
    }
}

static int is_method(const uir_task *t) { return t && t->kind != K_PROJECTOR; }

/* ------------------------------------------------------------------------ */
/* construction                                                             */
/* ------------------------------------------------------------------------ */

static uir_task *task_alloc(task_kind k)
{
    uir_task *t = (uir_task *)calloc(1, sizeof *t);
    if (!t) {
        uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
        return NULL;
    }
    t->refcount = 1;
    t->kind = k;
    /* projector defaults, ufo-ir-projector-task.c:101-110 and
     * ufo-ir-parallel-projector-task.c:149-154 */
    t->axis_position = -1.0f;
    t->step = 0.0f;
    t->relaxation = 1.0f;
    t->correction_scale = 1.0f;
    t->is_forward = 1;
    strcpy(t->model, "joseph");
    t->first_run = 1;
    /* method defaults */
    t->num_iterations = 10;          /* ufo-ir-method-task.c:109 */
    t->relaxation_factor = 0.25f;    /* ufo-ir-sirt-task.c:89, ufo-ir-sart-task.c:93 */
    t->positive_constraint = (k == K_ASDPOCS);
    t->beta = 1.0f;                  /* ufo-ir-asdpocs-task.c:173-179 */
    t->beta_red = 0.995f;
    t->ng = 20;
    t->alpha = 0.2f;
    t->alpha_red = 0.95f;
    t->r_max = 0.95f;
    t->lambda = 0.1f;                /* ufo-ir-sbtv-task.c (class_init defaults) */
    t->mu = 0.5f;
    t->compat = 1;
    return t;
}

uir_task *uir_ir_parallel_projector_task_new(void) { return task_alloc(K_PROJECTOR); }
uir_task *uir_ir_sirt_task_new(void) { return task_alloc(K_SIRT); }
uir_task *uir_ir_sart_task_new(void) { return task_alloc(K_SART); }
uir_task *uir_ir_asdpocs_task_new(void) { return task_alloc(K_ASDPOCS); }
uir_task *uir_ir_sbtv_task_new(void) { return task_alloc(K_SBTV); }

static void canon(const char *in, char *out, size_t n)
{
    size_t i = 0;
    for (; in[i] && i + 1 < n; ++i)
        out[i] = in[i] == '_' ? '-' : (char)tolower((unsigned char)in[i]);
    out[i] = 0;
}

uir_task *uir_task_new(const char *plugin)
{
    if (!plugin) {
        uir_set_error(UIR_ERR_INVALID_ARGUMENT, "plugin is NULL");
        return NULL;
    }
    char c[64];
    canon(plugin, c, sizeof c);
    for (int k = 0; k < 5; ++k)
        if (strcmp(c, g_plugin_names[k]) == 0)
            return task_alloc((task_kind)k);
    uir_set_error(UIR_ERR_NOT_FOUND, "no plugin '%s' in package ir", plugin);
    return NULL;
}

uir_task *uir_task_ref(uir_task *t)
{
    if (t)
        __atomic_add_fetch(&t->refcount, 1, __ATOMIC_SEQ_CST);
    return t;
}

static void task_free_tmp(uir_task *t)
{
    for (int i = 0; i < N_TMP; ++i) {
        uir_dfree(t->tmp[i]);
        t->tmp[i] = NULL;
        t->tmp_bytes[i] = 0;
    }
}

void uir_task_unref(uir_task *t)
{
    if (!t)
        return;
    if (__atomic_sub_fetch(&t->refcount, 1, __ATOMIC_SEQ_CST) > 0)
        return;
    if (t->res)
        (void)hipStreamSynchronize(t->res->stream);
    task_free_tmp(t);
    if (t->sart_exec)
        (void)hipGraphExecDestroy(t->sart_exec);
    if (t->sart_graph)
        (void)hipGraphDestroy(t->sart_graph);
    uir_resources_free(t->res);
    free(t->h_sin);
    free(t->h_cos);
    uir_dfree(t->d_sin);
    uir_dfree(t->d_cos);
    free(t->runs);
    uir_dfree(t->d_fp_items);
    uir_dfree(t->d_bp_items);
    uir_dfree(t->d_trig);
    uir_dfree(t->d_err);
    free(t->h_line);
    uir_dfree(t->d_line);
    free(t->line_runs);
    uir_dfree(t->d_ltrig);
    uir_dfree(t->d_lk);
    uir_dfree(t->d_lfp_items);
    uir_dfree(t->d_lbp_items);
    free(t->diag);
    uir_task_unref(t->projector);
    uir_task_unref(t->df_minimizer);
    free(t);
}

const char *uir_task_plugin(const uir_task *t) { return t ? g_plugin_names[t->kind] : NULL; }

/* ------------------------------------------------------------------------ */
/* properties                                                               */
/* ------------------------------------------------------------------------ */

static const prop_def *find_prop(const uir_task *t, const char *name, prop_type want)
{
    if (!t || !name) {
        uir_set_error(UIR_ERR_INVALID_ARGUMENT, "task or property name is NULL");
        return NULL;
    }
    char c[64];
    canon(name, c, sizeof c);
    for (const prop_def *p = props_of(t->kind); p->name; ++p) {
        if (strcmp(p->name, c) == 0) {
            if (p->type != want) {
                uir_set_error(UIR_ERR_INVALID_ARGUMENT, "property '%s' of '%s' has another type", name,
                              g_plugin_names[t->kind]);
                return NULL;
            }
            return p;
        }
    }
    uir_set_error(UIR_ERR_NOT_FOUND, "'%s' has no property '%s'", g_plugin_names[t->kind], name);
    return NULL;
}

static uir_status prop_status(void)
{
    const char *e = uir_last_error();
    return strstr(e, "has no property") ? UIR_ERR_NOT_FOUND : UIR_ERR_INVALID_ARGUMENT;
}

uir_status uir_task_set_float(uir_task *t, const char *name, float v)
{
    const prop_def *p = find_prop(t, name, P_FLOAT);
    if (!p)
        return prop_status();
    if (!(v >= p->min && v <= p->max))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "value %g is out of range [%g, %g] for property '%s'",
                             (double)v, p->min, p->max, name);
    *(float *)((char *)t + p->offset) = v;
    return UIR_OK;
}

uir_status uir_task_set_uint(uir_task *t, const char *name, uint32_t v)
{
    const prop_def *p = find_prop(t, name, P_UINT);
    if (!p)
        return prop_status();
    if (!((double)v >= p->min && (double)v <= p->max))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "value %u is out of range [%g, %g] for property '%s'",
                             v, p->min, p->max, name);
    *(uint32_t *)((char *)t + p->offset) = v;
    return UIR_OK;
}

uir_status uir_task_set_bool(uir_task *t, const char *name, int v)
{
    const prop_def *p = find_prop(t, name, P_BOOL);
    if (!p)
        return prop_status();
    *(int *)((char *)t + p->offset) = v ? 1 : 0;
    return UIR_OK;
}

uir_status uir_task_set_string(uir_task *t, const char *name, const char *v)
{
    const prop_def *p = find_prop(t, name, P_STRING);
    if (!p)
        return prop_status();
    if (!v)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL string");
    char *dst = (char *)t + p->offset;
    size_t i = 0;
    for (; v[i] && i + 1 < sizeof t->model; ++i) /* g_ascii_strdown, :283-287 */
        dst[i] = (char)tolower((unsigned char)v[i]);
    dst[i] = 0;
    return UIR_OK;
}

uir_status uir_task_set_object(uir_task *t, const char *name, uir_task *v)
{
    const prop_def *p = find_prop(t, name, P_OBJECT);
    if (!p)
        return prop_status();
    if (p->offset == OFF(projector) && v && v->kind != K_PROJECTOR)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "projector must be a projector task");
    if (p->offset == OFF(df_minimizer) && v && !is_method(v))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "df_minimizer must be a method task");
    uir_task **slot = (uir_task **)((char *)t + p->offset);
    uir_task_ref(v);
    uir_task_unref(*slot);
    *slot = v;
    return UIR_OK;
}

uir_status uir_task_property_type(const uir_task *t, const char *name, uir_prop_type *type)
{
    if (!t || !name || !type)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    char c[64];
    canon(name, c, sizeof c);
    for (const prop_def *p = props_of(t->kind); p->name; ++p)
        if (strcmp(p->name, c) == 0) {
            *type = (uir_prop_type)p->type;
            return UIR_OK;
        }
    return uir_set_error(UIR_ERR_NOT_FOUND, "'%s' has no property '%s'", g_plugin_names[t->kind], name);
}

uir_status uir_task_get_float(const uir_task *t, const char *name, float *v)
{
    const prop_def *p = find_prop(t, name, P_FLOAT);
    if (!p)
        return prop_status();
    *v = *(const float *)((const char *)t + p->offset);
    return UIR_OK;
}

uir_status uir_task_get_uint(const uir_task *t, const char *name, uint32_t *v)
{
    const prop_def *p = find_prop(t, name, P_UINT);
    if (!p)
        return prop_status();
    *v = *(const uint32_t *)((const char *)t + p->offset);
    return UIR_OK;
}

uir_status uir_task_get_bool(const uir_task *t, const char *name, int *v)
{
    const prop_def *p = find_prop(t, name, P_BOOL);
    if (!p)
        return prop_status();
    *v = *(const int *)((const char *)t + p->offset);
    return UIR_OK;
}

uir_status uir_task_get_string(const uir_task *t, const char *name, char *buf, size_t len)
{
    const prop_def *p = find_prop(t, name, P_STRING);
    if (!p)
        return prop_status();
    if (!buf || len == 0)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "empty buffer");
    snprintf(buf, len, "%s", (const char *)t + p->offset);
    return UIR_OK;
}

uir_status uir_task_get_object(const uir_task *t, const char *name, uir_task **v)
{
    const prop_def *p = find_prop(t, name, P_OBJECT);
    if (!p)
        return prop_status();
    *v = *(uir_task *const *)((const char *)t + p->offset);
    return UIR_OK;
}

/* ------------------------------------------------------------------------ */
/* temporaries                                                              */
/* ------------------------------------------------------------------------ */

static uir_status tmp_get(uir_task *t, int slot, size_t bytes, float **out)
{
    if (t->tmp_bytes[slot] < bytes) {
        uir_dfree(t->tmp[slot]);
        t->tmp[slot] = NULL;
        t->tmp_bytes[slot] = 0;
        uir_status s = uir_dmalloc(&t->tmp[slot], bytes);
        if (s != UIR_OK)
            return s;
        t->tmp_bytes[slot] = bytes;
    }
    *out = (float *)t->tmp[slot];
    return UIR_OK;
}

#define TRY(x) do { uir_status s_ = (x); if (s_ != UIR_OK) return s_; } while (0)
#define HIPTRY(x, what) TRY(uir_check_hip((x), (what)))

/* ------------------------------------------------------------------------ */
/* setup                                                                    */
/* ------------------------------------------------------------------------ */

uir_status uir_task_setup(uir_task *t, uir_resources *res)
{
    if (!t || !res)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "task or resources is NULL");
    TRY(uir_check_hip(hipSetDevice(res->device), "hipSetDevice"));
    if (t->res != res) {
        uir_resources_retain(res);
        uir_resources_free(t->res);
        t->res = res;
    }
    if (t->kind == K_PROJECTOR) {
        /* ufo-ir-parallel-projector-task.c:360-388 loads projector-parallel-<model>.cl
         * and fails for a model without kernels; the reference ships joseph,
         * this library also the line model (uir_line.hip) */
        if (strcmp(t->model, "joseph") == 0)
            t->line = 0;
        else if (strcmp(t->model, "line") == 0)
            t->line = 1;
        else
            return uir_set_error(UIR_ERR_SETUP, "no projection model '%s' (joseph or line)", t->model);
        return UIR_OK;
    }
    if (!t->projector) /* ufo-ir-method-task.c:212-215 */
        return uir_set_error(UIR_ERR_SETUP, "No projector specified.");
    TRY(uir_task_setup(t->projector, res));
    if (t->kind == K_ASDPOCS) {
        if (!t->df_minimizer) /* ufo-ir-asdpocs-task.c:389-392 */
            return uir_set_error(UIR_ERR_SETUP, "df_minimizer is not defined");
        if (t->df_minimizer->kind != K_SART && t->df_minimizer->kind != K_SIRT)
            return uir_set_error(UIR_ERR_SETUP, "df_minimizer must be sart or sirt");
        TRY(uir_task_set_object(t->df_minimizer, "projector", t->projector)); /* :395 */
        TRY(uir_task_setup(t->df_minimizer, res));
    }
    return UIR_OK;
}

/* ------------------------------------------------------------------------ */
/* projector geometry (get_requisition, ufo-ir-parallel-projector-task.c:305-347) */
/* ------------------------------------------------------------------------ */

static uir_status proj_build_luts(uir_task *p, uint32_t n_angles)
{
    free(p->h_sin);
    free(p->h_cos);
    free(p->runs);
    p->h_sin = (float *)malloc(sizeof(float) * (n_angles ? n_angles : 1));
    p->h_cos = (float *)malloc(sizeof(float) * (n_angles ? n_angles : 1));
    p->runs = (uir_subset *)malloc(sizeof(uir_subset) * (n_angles ? n_angles : 1));
    if (!p->h_sin || !p->h_cos || !p->runs)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
    uir_geom_lut(n_angles, p->step, p->h_sin, p->h_cos);
    p->n_runs = uir_geom_runs(p->h_sin, p->h_cos, n_angles, p->runs);
    uir_dfree(p->d_sin);
    uir_dfree(p->d_cos);
    p->d_sin = p->d_cos = NULL;
    TRY(uir_dmalloc((void **)&p->d_sin, sizeof(float) * (n_angles ? n_angles : 1)));
    TRY(uir_dmalloc((void **)&p->d_cos, sizeof(float) * (n_angles ? n_angles : 1)));
    if (n_angles) {
        hipStream_t st = p->res ? p->res->stream : 0;
        HIPTRY(hipMemcpyAsync(p->d_sin, p->h_sin, sizeof(float) * n_angles, hipMemcpyHostToDevice, st), "LUT upload");
        HIPTRY(hipMemcpyAsync(p->d_cos, p->h_cos, sizeof(float) * n_angles, hipMemcpyHostToDevice, st), "LUT upload");
        HIPTRY(hipStreamSynchronize(st), "LUT upload sync");
    }
    p->lut_n = n_angles;
    p->lut_gen++;
    return UIR_OK;
}

static uir_status proj_requisition(uir_task *p, const uir_buffer *in, uir_requisition *out)
{
    if (!in || in->dims[0] == 0 || in->dims[1] == 0)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "empty input buffer");
    if (p->angles_num == 0)
        p->angles_num = in->dims[1];
    if (p->first_run || p->angles_num != in->dims[1] || p->lut_n != p->angles_num) {
        p->first_run = 0;
        TRY(proj_build_luts(p, p->angles_num));
    }
    if (p->is_forward) {
        /* FP of an N x N image: n_dets from detectors-num (reference: input width). */
        p->n_dets = p->detectors_num ? p->detectors_num : in->dims[0];
        out->n_dims = 2;
        out->dims[0] = p->n_dets;
        out->dims[1] = p->angles_num;
    } else {
        p->n_dets = in->dims[0];
        uint32_t n = p->image_size ? p->image_size : p->n_dets;
        out->n_dims = 2;
        out->dims[0] = n;
        out->dims[1] = n;
    }
    return UIR_OK;
}

uir_status uir_task_get_requisition(uir_task *t, const uir_buffer *inputs, uir_requisition *out)
{
    if (!t || !inputs || !out)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (t->kind == K_PROJECTOR)
        return proj_requisition(t, &inputs[0], out);
    if (!t->projector) /* ufo-ir-method-task.c:247-251 */
        return uir_set_error(UIR_ERR_SETUP, "Projector not specified");
    return proj_requisition(t->projector, &inputs[0], out);
}

/* image edge of the projector's current geometry */
static uint32_t proj_image_n(const uir_task *p, uint32_t fallback)
{
    return p->image_size ? p->image_size : fallback;
}

static float proj_axis(const uir_task *p) { return uir_geom_axis(p->axis_position, p->n_dets); }

static void proj_runs_dev(const uir_task *p, const uir_subset *subs, uint32_t n, uirk_runs *r, uint32_t first)
{
    memset(r, 0, sizeof *r);
    uint32_t k = 0;
    for (uint32_t i = first; i < n && k < UIRK_MAX_RUNS; ++i, ++k) {
        r->offset[k] = subs[i].offset;
        r->n[k] = subs[i].n;
    }
    r->count = (int32_t)k;
    (void)p;
}

/* Work tables of the projector's current geometry for an n x n image. */
static uir_status proj_plan(uir_task *p, uint32_t n)
{
    const float axis = proj_axis(p);
    if ((uint64_t)n * n * 4 >= (1ull << 32) || (uint64_t)p->lut_n * p->n_dets * 4 >= (1ull << 32))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "image or sinogram of 4 GiB or more");
    if (p->d_fp_items && p->plan_n == n && p->plan_dets == p->n_dets && p->plan_angles == p->lut_n &&
        p->plan_axis == axis)
        return UIR_OK;
    uir_dfree(p->d_fp_items);
    uir_dfree(p->d_bp_items);
    uir_dfree(p->d_trig);
    p->d_fp_items = NULL;
    p->d_bp_items = NULL;
    p->d_trig = NULL;
    uirk_fp_vec fv = {0};
    uirk_bp_vec bv = {0};
    float *trig = (float *)malloc(sizeof(float) * 2 * (p->lut_n ? p->lut_n : 1));
    uir_status s = UIR_OK;
    if (!trig || uirk_plan_fp(&fv, 0, (int)n, (int)p->n_dets, axis, p->h_sin, p->h_cos, p->runs, p->n_runs,
                              UIRK_FP_ANGLES, UIRK_FP_WMAX_SINGLE) ||
        uirk_plan_bp(&bv, 0, (int)n, UIRK_BP_TILE)) {
        s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "work table");
        goto done;
    }
    uirk_plan_trig(p->h_sin, p->h_cos, p->lut_n, trig);
    if ((s = uir_dmalloc((void **)&p->d_fp_items, fv.n * sizeof(uirk_fp_item))) != UIR_OK ||
        (s = uir_dmalloc((void **)&p->d_bp_items, bv.n * sizeof(uirk_bp_item))) != UIR_OK ||
        (s = uir_dmalloc((void **)&p->d_trig, 2 * sizeof(float) * (p->lut_n ? p->lut_n : 1))) != UIR_OK)
        goto done;
    if (!p->d_err && (s = uir_dmalloc((void **)&p->d_err, sizeof(int))) != UIR_OK)
        goto done;
    if ((s = uir_check_hip(hipMemcpy(p->d_fp_items, fv.v, fv.n * sizeof(uirk_fp_item), hipMemcpyHostToDevice),
                           "table upload")) != UIR_OK ||
        (s = uir_check_hip(hipMemcpy(p->d_bp_items, bv.v, bv.n * sizeof(uirk_bp_item), hipMemcpyHostToDevice),
                           "table upload")) != UIR_OK ||
        (s = uir_check_hip(hipMemcpy(p->d_trig, trig, 2 * sizeof(float) * p->lut_n, hipMemcpyHostToDevice),
                           "table upload")) != UIR_OK ||
        (s = uir_check_hip(hipMemset(p->d_err, 0, sizeof(int)), "memset")) != UIR_OK)
        goto done;
    p->n_fp_items = (int)fv.n;
    p->n_bp_items = (int)bv.n;
    {
        double *cost = (double *)malloc(sizeof(double) * (fv.n ? fv.n : 1));
        if (!cost) {
            s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
            goto done;
        }
        for (size_t i = 0; i < fv.n; ++i)
            cost[i] = uirk_fp_item_cost(&fv.v[i]);
        uirk_plan_xcd(cost, (int)fv.n, uirk_plan_num_xcd(), &p->fp_part);
        uirk_plan_xcd(NULL, (int)bv.n, uirk_plan_num_xcd(), &p->bp_part);
        free(cost);
    }
    p->plan_n = n;
    p->plan_dets = p->n_dets;
    p->plan_angles = p->lut_n;
    p->plan_axis = axis;
done:
    free(fv.v);
    free(bv.v);
    free(trig);
    if (s != UIR_OK) {
        uir_dfree(p->d_fp_items);
        p->d_fp_items = NULL;
    }
    return s;
}

/* The line model's constants for an n x n image in the current geometry:
 * the per-angle table of the single-angle kernels (uir_line.hip), the runs
 * of angles of one family, and the fast path's trig pairs, constants and
 * work tables (mode 2 of the table kernels, for projections over all
 * angles). */
/* BP tiles of the task path's line mode: 64 x 32 pixels, twice as many
 * workgroups as 64 x 64 tiles (one slice per launch leaves few). */
#define LINE_BP_ROWS 32
static const uirk_tile g_line_tile = {UIRK_FP_ANGLES, 32, LINE_BP_ROWS, 0};

static uir_status line_plan(uir_task *p, uint32_t n)
{
    const float axis = proj_axis(p);
    if ((uint64_t)n * n * 4 >= (1ull << 32) || (uint64_t)p->lut_n * p->n_dets * 4 >= (1ull << 32))
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "image or sinogram of 4 GiB or more");
    if (n > 16384)
        return uir_set_error(UIR_ERR_UNSUPPORTED, "the line model takes images of at most 16384 x 16384");
    if (p->lut_n > 65535)
        return uir_set_error(UIR_ERR_UNSUPPORTED, "the line model takes at most 65535 angles");
    if (p->d_line && p->line_gen == p->lut_gen && p->line_n == n && p->line_angles == p->lut_n &&
        p->line_axis == axis)
        return UIR_OK;
    const uint32_t A = p->lut_n ? p->lut_n : 1;
    uirk_line_angle *h = (uirk_line_angle *)realloc(p->h_line, sizeof *h * A);
    if (!h)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
    p->h_line = h;
    uir_subset *runs = (uir_subset *)realloc(p->line_runs, sizeof *runs * A);
    if (!runs)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
    p->line_runs = runs;
    float *trig = (float *)malloc(sizeof(float) * 2 * A), *lk = (float *)malloc(sizeof(float) * 8 * A);
    float *su = (float *)malloc(sizeof(float) * A), *cu = (float *)malloc(sizeof(float) * A);
    uirk_fp_vec fv = {0};
    uirk_bp_vec bv = {0};
    double *cost = NULL;
    uir_status st = UIR_OK;
    if (!trig || !lk || !su || !cu) {
        st = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
        goto done;
    }
    uir_line_table(p->h_sin, p->h_cos, p->lut_n, n, axis, h, trig, lk);
    /* runs of one family; as work-table subsets their direction is joseph's
     * (1 = a sample per row, the line model's family 0) */
    uint32_t k = 0;
    for (uint32_t a = 0; a < p->lut_n; ++a) {
        if (k && runs[k - 1].direction == !h[a].family) {
            runs[k - 1].n++;
        } else {
            runs[k].offset = a;
            runs[k].n = 1;
            runs[k].direction = !h[a].family;
            ++k;
        }
    }
    p->n_line_runs = k;
    for (uint32_t a = 0; a < p->lut_n; ++a) {
        su[a] = trig[2 * a];
        cu[a] = trig[2 * a + 1];
    }
    if (uirk_plan_fp_line(&fv, 0, (int)n, (int)p->n_dets, axis, su, cu, runs, k, UIRK_FP_ANGLES, UIRK_FP_WMAX) ||
        uirk_plan_bp(&bv, 0, (int)n, LINE_BP_ROWS)) {
        st = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "work table");
        goto done;
    }
    uir_dfree(p->d_line);
    uir_dfree(p->d_ltrig);
    uir_dfree(p->d_lk);
    uir_dfree(p->d_lfp_items);
    uir_dfree(p->d_lbp_items);
    p->d_line = NULL;
    p->d_ltrig = p->d_lk = NULL;
    p->d_lfp_items = NULL;
    p->d_lbp_items = NULL;
    if ((st = uir_dmalloc((void **)&p->d_line, sizeof *h * A)) != UIR_OK ||
        (st = uir_dmalloc((void **)&p->d_ltrig, sizeof(float) * 2 * A)) != UIR_OK ||
        (st = uir_dmalloc((void **)&p->d_lk, sizeof(float) * 8 * A)) != UIR_OK ||
        (st = uir_dmalloc((void **)&p->d_lfp_items, (fv.n ? fv.n : 1) * sizeof(uirk_fp_item))) != UIR_OK ||
        (st = uir_dmalloc((void **)&p->d_lbp_items, (bv.n ? bv.n : 1) * sizeof(uirk_bp_item))) != UIR_OK)
        goto done;
    if (!p->d_err) {
        if ((st = uir_dmalloc((void **)&p->d_err, sizeof(int))) != UIR_OK ||
            (st = uir_check_hip(hipMemset(p->d_err, 0, sizeof(int)), "memset")) != UIR_OK)
            goto done;
    }
    if ((st = uir_check_hip(hipMemcpy(p->d_line, h, sizeof *h * p->lut_n, hipMemcpyHostToDevice), "table upload")) !=
            UIR_OK ||
        (st = uir_check_hip(hipMemcpy(p->d_ltrig, trig, sizeof(float) * 2 * p->lut_n, hipMemcpyHostToDevice),
                            "table upload")) != UIR_OK ||
        (st = uir_check_hip(hipMemcpy(p->d_lk, lk, sizeof(float) * 8 * p->lut_n, hipMemcpyHostToDevice),
                            "table upload")) != UIR_OK ||
        (st = uir_check_hip(hipMemcpy(p->d_lfp_items, fv.v, fv.n * sizeof(uirk_fp_item), hipMemcpyHostToDevice),
                            "table upload")) != UIR_OK ||
        (st = uir_check_hip(hipMemcpy(p->d_lbp_items, bv.v, bv.n * sizeof(uirk_bp_item), hipMemcpyHostToDevice),
                            "table upload")) != UIR_OK)
        goto done;
    p->n_lfp_items = (int)fv.n;
    p->n_lbp_items = (int)bv.n;
    cost = (double *)malloc(sizeof(double) * (fv.n ? fv.n : 1));
    if (!cost) {
        st = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
        goto done;
    }
    for (size_t i = 0; i < fv.n; ++i)
        cost[i] = uirk_fp_item_cost(&fv.v[i]);
    uirk_plan_xcd(cost, (int)fv.n, uirk_plan_num_xcd(), &p->lfp_part);
    uirk_plan_xcd(NULL, (int)bv.n, uirk_plan_num_xcd(), &p->lbp_part);
    p->line_gen = p->lut_gen;
    p->line_n = n;
    p->line_angles = p->lut_n;
    p->line_axis = axis;
done:
    free(trig);
    free(lk);
    free(su);
    free(cu);
    free(fv.v);
    free(bv.v);
    free(cost);
    if (st != UIR_OK) {
        uir_dfree(p->d_line);
        p->d_line = NULL; /* rebuilt on the next call */
    }
    return st;
}

/* A slice descriptor of the fast path's line mode for an n x n image. */
static void line_slice(const uir_task *p, uint32_t n, uirk_slice *sl)
{
    memset(sl, 0, sizeof *sl);
    sl->trig = p->d_ltrig;
    sl->line = p->d_lk;
    sl->n = (int32_t)n;
    sl->n_dets = (int32_t)p->n_dets;
    sl->n_angles = (int32_t)p->lut_n;
    sl->axis = proj_axis(p);
    sl->corr = p->correction_scale;
    sl->relax = p->relaxation;
    sl->pix_scale = 1.0f;
    sl->n_runs = 1;
    sl->run_off[0] = 0;
    sl->run_n[0] = p->lut_n;
}

/* Line-model FP of angles [a0, a0 + na): over all angles one launch of the
 * fast path's table kernel, else one launch of the single-angle kernels per
 * run of one family. */
static uir_status line_fp(uir_task *p, uint32_t n, const float *vol, const float *r_sino, float *w_sino,
                          const float *ray_w, uint32_t a0, uint32_t na, int few_rays, hipStream_t st)
{
    TRY(line_plan(p, n));
    if (a0 == 0 && na == p->lut_n && !few_rays) {
        uirk_slice sl;
        line_slice(p, n, &sl);
        sl.x = vol;
        sl.b = r_sino;
        sl.r = w_sino;
        sl.ray_w = ray_w;
        return uir_check_hip(uirk_fp_table(NULL, &sl, p->d_lfp_items, p->n_lfp_items, &p->lfp_part, 2, UIR_DTYPE_F32,
                                           UIR_DTYPE_F32, NULL, uirk_plan_grid_xcd(&p->lfp_part, p->n_lfp_items),
                                           p->d_err, st),
                             "FP launch");
    }
    for (uint32_t a = a0; a < a0 + na;) {
        const int fam = p->h_line[a].family;
        uint32_t e = a + 1;
        while (e < a0 + na && p->h_line[e].family == fam)
            ++e;
        HIPTRY(uirk_fp_line(vol, r_sino, w_sino, ray_w, p->d_line, fam, (int)n, (int)p->n_dets, (int)a,
                            (int)(e - a), p->correction_scale, few_rays, st), "FP launch");
        a = e;
    }
    return UIR_OK;
}

static void proj_slice(const uir_task *p, uint32_t n, uirk_slice *sl)
{
    memset(sl, 0, sizeof *sl);
    sl->trig = p->d_trig;
    sl->n = (int32_t)n;
    sl->n_dets = (int32_t)p->n_dets;
    sl->n_angles = (int32_t)p->lut_n;
    sl->axis = proj_axis(p);
    sl->corr = p->correction_scale;
    sl->relax = p->relaxation;
}

/* forward over all runs (ufo-ir-parallel-projector-task.c:390-408);
 * ray_w != NULL fuses the SIRT ray weighting. */
static uir_status proj_fp(uir_task *p, uint32_t n, const float *vol, const float *r_sino, float *w_sino,
                          const float *ray_w)
{
    if (p->line)
        return line_fp(p, n, vol, r_sino, w_sino, ray_w, 0, p->lut_n, 0, p->res->stream);
    TRY(proj_plan(p, n));
    uirk_slice sl;
    proj_slice(p, n, &sl);
    sl.x = vol;
    sl.b = r_sino;
    sl.r = w_sino;
    sl.ray_w = ray_w;
    return uir_check_hip(uirk_fp_table(NULL, &sl, p->d_fp_items, p->n_fp_items, &p->fp_part, 0, UIR_DTYPE_F32, UIR_DTYPE_F32, NULL,
                                       uirk_plan_grid_xcd(&p->fp_part, p->n_fp_items), p->d_err, p->res->stream), "FP launch");
}

/* backward over all runs (:410-429).  pix_w == NULL: accumulate form,
 * vol += relax * BP per run; else the fused SIRT update with POSC option. */
static uir_status proj_bp_ex(uir_task *p, uint32_t n, const float *sino, float *vol, const float *pix_w, int posc)
{
    if (p->line) { /* the fast path's table kernel in line mode */
        TRY(line_plan(p, n));
        uirk_slice sl;
        line_slice(p, n, &sl);
        sl.b = sino;
        sl.x = vol;
        sl.x_out = vol;
        sl.pix_w = pix_w;
        sl.posc = posc;
        return uir_check_hip(uirk_bp_table(NULL, &sl, p->d_lbp_items, p->n_lbp_items, &p->lbp_part, 2, UIR_DTYPE_F32,
                                           UIR_DTYPE_F32, &g_line_tile, uirk_plan_grid_xcd(&p->lbp_part, p->n_lbp_items),
                                           p->d_err, p->res->stream),
                             "BP launch");
    }
    TRY(proj_plan(p, n));
    if (pix_w && p->n_runs > UIRK_MAX_RUNS)
        return uir_set_error(UIR_ERR_UNSUPPORTED, "more than %d direction runs", UIRK_MAX_RUNS);
    for (uint32_t first = 0; first < p->n_runs; first += UIRK_MAX_RUNS) {
        uirk_slice sl;
        proj_slice(p, n, &sl);
        sl.b = sino;
        sl.x = vol;
        sl.x_out = vol;
        sl.pix_w = pix_w;
        sl.posc = posc;
        uint32_t k = 0;
        for (uint32_t i = first; i < p->n_runs && k < UIRK_MAX_RUNS; ++i, ++k) {
            sl.run_off[k] = p->runs[i].offset;
            sl.run_n[k] = p->runs[i].n;
        }
        sl.n_runs = (int32_t)k;
        TRY(uir_check_hip(uirk_bp_table(NULL, &sl, p->d_bp_items, p->n_bp_items, &p->bp_part, 0, UIR_DTYPE_F32, UIR_DTYPE_F32, NULL,
                                        uirk_plan_grid_xcd(&p->bp_part, p->n_bp_items), p->d_err, p->res->stream), "BP launch"));
    }
    return UIR_OK;
}

static uir_status proj_bp(uir_task *p, uint32_t n, const float *sino, float *vol)
{
    return proj_bp_ex(p, n, sino, vol, NULL, 0);
}

/* The staging-window overflow flag of the table kernels (never set when the
 * work tables come from uirk_plan_fp); synchronises the stream. */
static uir_status uir_check_err_flag(uir_task *p)
{
    if (!p->d_err)
        return UIR_OK;
    int e = 0;
    TRY(uir_check_hip(hipMemcpyAsync(&e, p->d_err, sizeof e, hipMemcpyDeviceToHost, p->res->stream), "flag read"));
    TRY(uir_check_hip(hipStreamSynchronize(p->res->stream), "hipStreamSynchronize"));
    if (e)
        return uir_set_error(UIR_ERR_DEVICE, "projector footprint exceeded the staging window (flag %d)", e);
    return UIR_OK;
}

static uir_status check_buf(const uir_buffer *b, uint64_t w, uint64_t h, const char *what)
{
    if (!b || !b->data)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: NULL buffer", what);
    if (b->dtype != UIR_DTYPE_F32)
        return uir_set_error(UIR_ERR_UNSUPPORTED, "%s: tasks take FP32 buffers", what);
    if (b->dims[0] != w || b->dims[1] != h)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: dims %ux%u, expected %llux%llu", what, b->dims[0],
                             b->dims[1], (unsigned long long)w, (unsigned long long)h);
    return UIR_OK;
}

static uir_status proj_ready(uir_task *p)
{
    if (!p || p->kind != K_PROJECTOR)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "not a projector task");
    if (!p->res)
        return uir_set_error(UIR_ERR_SETUP, "projector is not set up");
    if (!p->h_sin)
        return uir_set_error(UIR_ERR_SETUP, "projector geometry unknown: call get_requisition first");
    return UIR_OK;
}

uir_status uir_projector_forward(uir_task *p, const uir_buffer *in, uir_buffer *out)
{
    TRY(proj_ready(p));
    uint32_t n = in ? in->dims[0] : 0;
    TRY(check_buf(in, n, n, "volume"));
    TRY(check_buf(out, p->n_dets, p->lut_n, "sinogram"));
    return proj_fp(p, n, (const float *)in->data, (const float *)out->data, (float *)out->data, NULL);
}

uir_status uir_projector_backward(uir_task *p, const uir_buffer *in, uir_buffer *out)
{
    TRY(proj_ready(p));
    uint32_t n = out ? out->dims[0] : 0;
    TRY(check_buf(in, p->n_dets, p->lut_n, "sinogram"));
    TRY(check_buf(out, n, n, "volume"));
    return proj_bp(p, n, (const float *)in->data, (float *)out->data);
}

uir_status uir_parallel_projector_subset_fp(uir_task *p, const uir_buffer *volume, uir_buffer *sinogram,
                                            const uir_subset *sub)
{
    TRY(proj_ready(p));
    if (!sub || (uint64_t)sub->offset + sub->n > p->lut_n)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "subset outside the angle range");
    uint32_t n = volume ? volume->dims[0] : 0;
    TRY(check_buf(volume, n, n, "volume"));
    TRY(check_buf(sinogram, p->n_dets, p->lut_n, "sinogram"));
    if (p->line) /* each angle takes its own family; sub->direction is joseph's */
        return line_fp(p, n, (const float *)volume->data, (const float *)sinogram->data, (float *)sinogram->data,
                       NULL, sub->offset, sub->n, sub->n < 8, p->res->stream);
    return uir_check_hip(uirk_fp_ref((const float *)volume->data, (const float *)sinogram->data,
                                     (float *)sinogram->data, NULL, p->d_sin, p->d_cos, (int)n, (int)p->n_dets,
                                     proj_axis(p), p->correction_scale, (int)sub->offset, (int)sub->n,
                                     sub->direction, p->res->stream), "FP launch");
}

uir_status uir_parallel_projector_subset_bp(uir_task *p, uir_buffer *volume, const uir_buffer *sinogram,
                                            const uir_subset *sub)
{
    TRY(proj_ready(p));
    if (!sub || (uint64_t)sub->offset + sub->n > p->lut_n)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "subset outside the angle range");
    uint32_t n = volume ? volume->dims[0] : 0;
    TRY(check_buf(volume, n, n, "volume"));
    TRY(check_buf(sinogram, p->n_dets, p->lut_n, "sinogram"));
    if (p->line) {
        TRY(line_plan(p, n));
        return uir_check_hip(uirk_bp_line((const float *)volume->data, (float *)volume->data,
                                          (const float *)sinogram->data, NULL, 0, p->d_line, (int)n, (int)p->n_dets,
                                          (int)sub->offset, (int)sub->n, p->relaxation, p->res->stream),
                             "BP launch");
    }
    uirk_runs r;
    proj_runs_dev(p, sub, 1, &r, 0);
    return uir_check_hip(uirk_bp_ref((const float *)volume->data, (float *)volume->data,
                                     (const float *)sinogram->data, (int)p->lut_n, p->d_sin, p->d_cos, (int)n,
                                     (int)p->n_dets, proj_axis(p), p->relaxation, &r, NULL, 0, p->res->stream),
                         "BP launch");
}

uir_status uir_parallel_projector_get_luts(uir_task *p, const float **s, const float **c, uint32_t *n)
{
    if (!p || p->kind != K_PROJECTOR || !p->h_sin)
        return uir_set_error(UIR_ERR_SETUP, "projector geometry unknown");
    if (s) *s = p->h_sin;
    if (c) *c = p->h_cos;
    if (n) *n = p->lut_n;
    return UIR_OK;
}

uir_status uir_parallel_projector_get_subsets(uir_task *p, const uir_subset **subs, uint32_t *n)
{
    if (!p || p->kind != K_PROJECTOR || !p->runs)
        return uir_set_error(UIR_ERR_SETUP, "projector geometry unknown");
    if (subs) *subs = p->runs;
    if (n) *n = p->n_runs;
    return UIR_OK;
}

/* ------------------------------------------------------------------------ */
/* methods                                                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    uir_task *p;       /* projector */
    uint32_t n, D, A;  /* image edge, detectors, angles */
    size_t nv, ns;
    hipStream_t st;
} mctx;

static uir_status method_begin(uir_task *t, const uir_buffer *inputs, uir_buffer *out, mctx *c)
{
    if (!t->res)
        return uir_set_error(UIR_ERR_SETUP, "task is not set up");
    if (!t->projector)
        return uir_set_error(UIR_ERR_SETUP, "No projector specified.");
    if (!inputs)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "no input");
    uir_requisition req;
    TRY(uir_task_get_requisition(t, inputs, &req));
    c->p = t->projector;
    if (!c->p->res)
        c->p->res = uir_resources_retain(t->res);
    c->D = c->p->n_dets;
    c->A = c->p->lut_n;
    c->n = proj_image_n(c->p, c->D);
    c->nv = (size_t)c->n * c->n;
    c->ns = (size_t)c->A * c->D;
    c->st = t->res->stream;
    TRY(check_buf(&inputs[0], c->D, c->A, "sinogram"));
    TRY(check_buf(out, c->n, c->n, "output"));
    if (c->p->n_runs > UIRK_MAX_RUNS)
        return uir_set_error(UIR_ERR_UNSUPPORTED, "more than %d direction runs", UIRK_MAX_RUNS);
    return UIR_OK;
}

/* SIRT ray and pixel weights, ufo-ir-sirt-task.c:176-193. */
static uir_status sirt_weights(uir_task *t, mctx *c, float **ray_w, float **pix_w)
{
    float *vol_tmp, *sino_tmp;
    TRY(tmp_get(t, 0, c->nv * 4, &vol_tmp));
    TRY(tmp_get(t, 1, c->ns * 4, ray_w));
    TRY(tmp_get(t, 2, c->ns * 4, &sino_tmp));
    TRY(tmp_get(t, 3, c->nv * 4, pix_w));
    c->p->relaxation = 1.0f;
    c->p->correction_scale = 1.0f;
    HIPTRY(uirk_set(vol_tmp, c->nv, 1.0f, c->st), "set");
    HIPTRY(uirk_set(*ray_w, c->ns, 0.0f, c->st), "set");
    TRY(proj_fp(c->p, c->n, vol_tmp, *ray_w, *ray_w, NULL));
    HIPTRY(uirk_inv(*ray_w, *ray_w, c->ns, c->st), "inv");
    HIPTRY(uirk_set(sino_tmp, c->ns, 1.0f, c->st), "set");
    HIPTRY(uirk_set(*pix_w, c->nv, 0.0f, c->st), "set");
    TRY(proj_bp(c->p, c->n, sino_tmp, *pix_w));
    if (c->p->line) /* no pixel covered by less than UIRK_LINE_PIX_MIN of chord */
        HIPTRY(uirk_inv_min(*pix_w, *pix_w, c->nv, UIRK_LINE_PIX_MIN, c->st), "inv");
    else
        HIPTRY(uirk_inv(*pix_w, *pix_w, c->nv, c->st), "inv");
    return UIR_OK;
}

/* SIRT process, ufo-ir-sirt-task.c:163-222.  One iteration is two launches:
 * FP with the residual and ray weighting fused (b - A x) * R, and BP with the
 * pixel weighting and update fused x + C * relax * A^T r (+ POSC if enabled). */
static uir_status sirt_run(uir_task *t, const float *b, float *x, float relax, uint32_t iters, int posc,
                           int warm, mctx *c)
{
    float *ray_w, *pix_w, *sino_tmp;
    TRY(sirt_weights(t, c, &ray_w, &pix_w));
    TRY(tmp_get(t, 2, c->ns * 4, &sino_tmp));
    c->p->relaxation = relax;            /* :195-197 */
    c->p->correction_scale = -1.0f;
    if (!warm)
        HIPTRY(uirk_set(x, c->nv, 0.0f, c->st), "set");
    for (uint32_t it = 0; it < iters; ++it) { /* :202-215 */
        TRY(proj_fp(c->p, c->n, x, b, sino_tmp, ray_w));
        TRY(proj_bp_ex(c->p, c->n, sino_tmp, x, pix_w, posc));
    }
    return uir_check_err_flag(c->p);
}

/* SART process, ufo-ir-sart-task.c:173-228: per angle, FP of the row with the
 * ray weighting fused (subset_fp + op_mulRows), then BP of that row into x. */
/* One SART sweep (ufo-ir-sart-task.c:208-220): per angle in order, FP of that
 * angle with the residual and ray weights fused (w = (b - A_a x) R_a), then
 * its BP x = x + relax A_a^T w.  One angle is only n_dets rays, so the FP
 * gives each ray a wave, starting its 64-sample blocks from the geometry's
 * block-start table (starts; without one, 16 threads per ray replay the float
 * stepping); both are bit-identical to the one-thread-per-ray walk.  The BP
 * takes four pixels per thread. */
static uir_status sart_fp_angle(mctx *c, const float *x, float *sino_tmp, const float *ray_w, float axis,
                                const float *starts, uint32_t a)
{
    uir_task *p = c->p;
    if (p->line) /* the line model: each ray over 16 lanes */
        return uir_check_hip(uirk_fp_line(x, sino_tmp, sino_tmp, ray_w, p->d_line, p->h_line[a].family, (int)c->n,
                                          (int)c->D, (int)a, 1, p->correction_scale, 1, c->st), "FP launch");
    if (starts)
        return uir_check_hip(uirk_fp_ref_seg(x, sino_tmp, sino_tmp, ray_w, p->d_sin, p->d_cos, (int)c->n, (int)c->D,
                                             axis, p->correction_scale, (int)a, 1, -1, starts, c->st), "FP launch");
    return uir_check_hip(uirk_fp_ref_split(x, sino_tmp, sino_tmp, ray_w, p->d_sin, p->d_cos, (int)c->n, (int)c->D,
                                           axis, p->correction_scale, (int)a, 1, -1, c->st), "FP launch");
}

static uir_status sart_bp_angle(mctx *c, float *x, const float *sino_tmp, float axis, uint32_t a)
{
    uir_task *p = c->p;
    if (p->line)
        return uir_check_hip(uirk_bp_line(x, x, sino_tmp, NULL, 0, p->d_line, (int)c->n, (int)c->D, (int)a, 1,
                                          p->relaxation, c->st), "BP launch");
    return uir_check_hip(uirk_bp_ref_angle(x, x, sino_tmp, p->d_sin, p->d_cos, (int)c->n, (int)c->D, axis,
                                           p->relaxation, (int)a, c->st), "BP launch");
}

static uir_status sart_sweep(uir_task *t, mctx *c, const float *b, float *x, float *sino_tmp, const float *ray_w,
                             float axis, const float *starts)
{
    HIPTRY(hipMemcpyAsync(sino_tmp, b, c->ns * 4, hipMemcpyDeviceToDevice, c->st), "copy");
    for (uint32_t a = 0; a < c->A; ++a) {
        TRY(sart_fp_angle(c, x, sino_tmp, ray_w, axis, starts, a));
        TRY(sart_bp_angle(c, x, sino_tmp, axis, a));
    }
    (void)t;
    return UIR_OK;
}

/* Bytes of the FP block-start table of a geometry; tables above this bound
 * are not built (the FP then replays the stepping per thread). */
#define SART_STARTS_MAX_BYTES ((size_t)4 << 30)

/* SART set-up (ufo-ir-sart-task.c:193-207): ray weights R = inv(FP(1)) per
 * angle, corr = -1, relax = relaxation, x = 0 (unless warm); with `table`,
 * also the FP block-start table of the geometry. */
typedef struct {
    float *sino_tmp, *ray_w, *starts;
    float axis;
} sart_prep;

static uir_status sart_prepare(uir_task *t, float *x, float relax, int warm, int table, mctx *c, sart_prep *sp)
{
    float *vol_tmp;
    TRY(tmp_get(t, 4, c->ns * 4, &sp->sino_tmp));
    TRY(tmp_get(t, 5, c->nv * 4, &vol_tmp));
    TRY(tmp_get(t, 6, c->ns * 4, &sp->ray_w));
    HIPTRY(uirk_set(vol_tmp, c->nv, 1.0f, c->st), "set");          /* :193-200 */
    HIPTRY(uirk_set(sp->ray_w, c->ns, 0.0f, c->st), "set");
    c->p->correction_scale = 1.0f;
    TRY(proj_fp(c->p, c->n, vol_tmp, sp->ray_w, sp->ray_w, NULL));
    HIPTRY(uirk_inv(sp->ray_w, sp->ray_w, c->ns, c->st), "inv");
    c->p->correction_scale = -1.0f;                                 /* :205-207 */
    c->p->relaxation = relax;
    if (!warm)
        HIPTRY(uirk_set(x, c->nv, 0.0f, c->st), "set");
    sp->axis = proj_axis(c->p);
    sp->starts = NULL;
    /* the rays' block starts: the float stepping taken once per call */
    const size_t starts_bytes = (size_t)c->A * c->D * ((c->n + 63) / 64) * 4;
    if (c->p->line) /* the line model's table, built by the FP above */
        TRY(line_plan(c->p, c->n));
    else if (table && starts_bytes <= SART_STARTS_MAX_BYTES) {
        TRY(tmp_get(t, 25, starts_bytes, &sp->starts));
        HIPTRY(uirk_fp_block_starts(c->p->d_sin, c->p->d_cos, (int)c->n, (int)c->D, sp->axis, (int)c->A, -1,
                                    sp->starts, c->st), "block starts");
    }
    return UIR_OK;
}

static uir_status sart_run(uir_task *t, const float *b, float *x, float relax, uint32_t iters, int warm,
                           mctx *c)
{
    sart_prep sp;
    TRY(sart_prepare(t, x, relax, warm, iters > 0, c, &sp));
    if (iters == 0)
        return UIR_OK;
    float *sino_tmp = sp.sino_tmp, *ray_w = sp.ray_w, *starts = sp.starts;
    const float axis = sp.axis;
    /* 2 A launches per sweep: capture one sweep and replay it (:208-220) */
    const int same = t->sart_exec && t->sart_key.x == x && t->sart_key.b == b && t->sart_key.tmp == sino_tmp &&
                     t->sart_key.starts == starts && t->sart_key.line == (c->p->line ? c->p->d_line : NULL) &&
                     t->sart_key.ray_w == ray_w && t->sart_key.sin == c->p->d_sin && t->sart_key.cos == c->p->d_cos &&
                     t->sart_key.n == c->n && t->sart_key.D == c->D && t->sart_key.A == c->A &&
                     t->sart_key.axis == axis && t->sart_key.relax == c->p->relaxation &&
                     t->sart_key.corr == c->p->correction_scale;
    if (!same && iters == 1) /* a capture pays off from the second replay on */
        return sart_sweep(t, c, b, x, sino_tmp, ray_w, axis, starts);
    if (!same) {
        if (t->sart_exec)
            (void)hipGraphExecDestroy(t->sart_exec);
        if (t->sart_graph)
            (void)hipGraphDestroy(t->sart_graph);
        t->sart_exec = NULL;
        t->sart_graph = NULL;
        HIPTRY(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        uir_status s = sart_sweep(t, c, b, x, sino_tmp, ray_w, axis, starts);
        hipGraph_t g = NULL;
        const int e = hipStreamEndCapture(c->st, &g);
        if (s != UIR_OK || e != hipSuccess) {
            if (g)
                (void)hipGraphDestroy(g);
            return s != UIR_OK ? s : uir_check_hip(e, "hipStreamEndCapture");
        }
        t->sart_graph = g;
        HIPTRY(hipGraphInstantiate(&t->sart_exec, g, NULL, NULL, 0), "hipGraphInstantiate");
        t->sart_key.x = x;
        t->sart_key.b = b;
        t->sart_key.tmp = sino_tmp;
        t->sart_key.starts = starts;
        t->sart_key.line = c->p->line ? c->p->d_line : NULL;
        t->sart_key.ray_w = ray_w;
        t->sart_key.sin = c->p->d_sin;
        t->sart_key.cos = c->p->d_cos;
        t->sart_key.n = c->n;
        t->sart_key.D = c->D;
        t->sart_key.A = c->A;
        t->sart_key.axis = axis;
        t->sart_key.relax = c->p->relaxation;
        t->sart_key.corr = c->p->correction_scale;
    }
    for (uint32_t it = 0; it < iters; ++it)
        HIPTRY(hipGraphLaunch(t->sart_exec, c->st), "hipGraphLaunch");
    return UIR_OK;
}

/* warm != 0: the data step continues from x (textbook ASD-POCS) instead of zero. */
static uir_status method_process_df(uir_task *t, const float *b, float *x, mctx *c, float relax, int warm)
{
    if (t->kind == K_SIRT)
        return sirt_run(t, b, x, relax, t->num_iterations, t->positive_constraint, warm, c);
    return sart_run(t, b, x, relax, t->num_iterations, warm, c);
}

/* The textbook variants' distances: L2 norms, each element once, in FP64. */
static uir_status ws_l2(uir_workspace *ws, const float *v, size_t n, void *st, double *out)
{
    double d;
    TRY(uir_ws_dot_f64(ws, v, v, n, st, &d));
    *out = sqrt(d);
    return UIR_OK;
}

/* ASD-POCS process, ufo-ir-asdpocs-task.c:411-509 (tvstd :550-580).
 * compat = 0 runs ASD-POCS as published (Sidky and Pan, Phys. Med. Biol. 53
 * (2008) 4777): the data step continues from the current image, dd, dp and dg
 * are L2 norms with dp measured from the image at the start of the
 * iteration, the ng TV steps move dtgv along the L2-normalised gradient (a
 * zero gradient ends them), and the output is the image after TV. */
static uir_status asdpocs_run(uir_task *t, const float *b, float *out, mctx *c)
{
    uir_task *df = t->df_minimizer;
    if (!df)
        return uir_set_error(UIR_ERR_SETUP, "df_minimizer is not defined");
    if (df->projector != t->projector)
        TRY(uir_task_set_object(df, "projector", t->projector)); /* :434 */
    if (!df->res)
        df->res = uir_resources_retain(t->res);
    const int tb = !t->compat;
    float *x, *x_prev, *x_res, *b_res, *grad;
    TRY(tmp_get(t, 7, c->nv * 4, &x));
    TRY(tmp_get(t, 8, c->nv * 4, &x_prev));
    TRY(tmp_get(t, 9, c->nv * 4, &x_res));
    TRY(tmp_get(t, 10, c->ns * 4, &b_res));
    TRY(tmp_get(t, 11, c->nv * 4, &grad));
    HIPTRY(uirk_set(x, c->nv, 0.0f, c->st), "set");
    HIPTRY(uirk_set(x_prev, c->nv, 0.0f, c->st), "set");
    free(t->diag);
    t->diag = (float *)calloc(4 * (size_t)(t->num_iterations ? t->num_iterations : 1), sizeof(float));
    t->diag_rows = 0;
    float dp = 1.0f, dd = 1.0f, dg = 1.0f, dtgv = 1.0f;
    float beta = t->beta;
    const uint32_t ng = tb ? t->ng : 20; /* tvstd :573 takes 20 steps; ng is unused there */
    uir_workspace *ws = &t->res->ws;
#define NORM(v, n, out) (tb ? ws_l2(ws, (v), (n), c->st, (out)) : uir_ws_l1(ws, (v), (n), c->st, (out)))
    for (uint32_t it = 0; it < t->num_iterations; ++it) {
        /* g_object_set(df, "relaxation_factor", beta): GLib rejects values outside [0,1]. */
        if (beta >= 0.0f && beta <= 1.0f)
            df->relaxation_factor = beta;
        if (tb)
            HIPTRY(hipMemcpyAsync(x_prev, x, c->nv * 4, hipMemcpyDeviceToDevice, c->st), "copy");
        TRY(method_process_df(df, b, x, c, df->relaxation_factor, tb));    /* :460 */
        if (t->positive_constraint)                                         /* :463-465 */
            HIPTRY(uirk_unary(UIRK_POSC, x, x, c->nv, 0, c->st), "POSC");
        if (!tb)
            HIPTRY(hipMemcpyAsync(out, x, c->nv * 4, hipMemcpyDeviceToDevice, c->st), "copy"); /* :468 */
        c->p->correction_scale = -1.0f;                                     /* :471-475 */
        TRY(proj_fp(c->p, c->n, x, b, b_res, NULL));
        double d;
        TRY(NORM(b_res, c->ns, &d));                                        /* :478 */
        dd = (float)d;
        HIPTRY(uirk_binary(UIRK_SUB, x, x_prev, 0, x_res, c->nv, c->st), "deduction"); /* :481-482 */
        TRY(NORM(x_res, c->nv, &d));
        dp = (float)d;
        if (it == 0)                                                        /* :485-487 */
            dtgv = t->alpha * dp;
        HIPTRY(hipMemcpyAsync(x_prev, x, c->nv * 4, hipMemcpyDeviceToDevice, c->st), "copy"); /* :490 */
        HIPTRY(uirk_set(grad, c->nv, 0.0f, c->st), "set");                 /* tvstd :559 */
        for (uint32_t k = 0; k < ng; ++k) {                                 /* :573-579 */
            HIPTRY(uirk_l1_grad(x, grad, c->n, c->n, c->st), "l1_grad");
            TRY(NORM(grad, c->nv, &d));
            if (tb && !(d > 0))
                break;
            const float factor = dtgv / (float)d;
            HIPTRY(uirk_binary(UIRK_SUB_SCALED, x, grad, factor, x, c->nv, c->st), "deduction2");
        }
        HIPTRY(uirk_binary(UIRK_SUB, x, x_prev, 0, x_res, c->nv, c->st), "deduction"); /* :496-497 */
        TRY(NORM(x_res, c->nv, &d));
        dg = (float)d;
        beta *= t->beta_red;                                                /* :498 */
        t->diag[4 * it + 0] = dd;
        t->diag[4 * it + 1] = dp;
        t->diag[4 * it + 2] = dg;
        t->diag[4 * it + 3] = dtgv;
        t->diag_rows = it + 1;
        const float epsilon = 0.001f;                                       /* :495,501-503 */
        if (dg > t->r_max * dp && dd > epsilon)
            dtgv *= t->alpha_red;
        if (tb)
            HIPTRY(hipMemcpyAsync(out, x, c->nv * 4, hipMemcpyDeviceToDevice, c->st), "copy");
    }
#undef NORM
    return UIR_OK;
}

uir_status uir_asdpocs_get_diagnostics(uir_task *t, const float **rows4, uint32_t *n_rows)
{
    if (!t || t->kind != K_ASDPOCS)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "not an asdpocs task");
    if (rows4) *rows4 = t->diag;
    if (n_rows) *n_rows = t->diag_rows;
    return UIR_OK;
}

/* ---- SBTV (ufo-ir-sbtv-task.c) ---- */

typedef struct {
    uir_task *t;
    mctx *c;
    float *tempA, *tempAt, *tempD, *tempDxt;
} sbctx;

/* The CGS inner product: the reference's strided "dot"
 * (ufo-ir-basic-ops-processor.c:148-196) when compat, else every element once,
 * accumulated in FP64 and rounded to float. */
static uir_status sb_dot(sbctx *s, const float *a, const float *b, float *out)
{
    if (s->t->compat)
        return uir_ws_dot_compat(&s->t->res->ws, a, b, s->c->n, s->c->n, s->c->st, out);
    double d;
    TRY(uir_ws_dot_f64(&s->t->res->ws, a, b, s->c->nv, s->c->st, &d));
    *out = (float)d;
    return UIR_OK;
}

static uir_status sb_l2(sbctx *s, const float *v, float *out)
{
    float d;
    TRY(sb_dot(s, v, v, &d));
    *out = (float)sqrt(d);
    return UIR_OK;
}

/* processA, :587-635 */
static uir_status sb_process_a(sbctx *s, const float *in, float *out)
{
    mctx *c = s->c;
    const uint32_t n = c->n;
    HIPTRY(uirk_set(out, c->nv, 0.0f, c->st), "set");
    HIPTRY(uirk_set(s->tempA, c->ns, 0.0f, c->st), "set");
    HIPTRY(uirk_set(s->tempAt, c->nv, 0.0f, c->st), "set");
    TRY(proj_fp(c->p, n, in, s->tempA, s->tempA, NULL));
    TRY(proj_bp(c->p, n, s->tempA, s->tempAt));
    HIPTRY(uirk_unary(UIRK_SCALE, s->tempAt, s->tempAt, c->nv, s->t->mu, c->st), "mul_scalar");
    HIPTRY(uirk_diff(UIRK_DY, in, s->tempD, n, n, c->st), "Dy");
    HIPTRY(uirk_diff(UIRK_DYT, s->tempD, out, n, n, c->st), "Dyt");
    HIPTRY(uirk_diff(UIRK_DX, in, s->tempD, n, n, c->st), "Dx");
    HIPTRY(uirk_diff(UIRK_DXT, s->tempD, s->tempDxt, n, n, c->st), "Dxt");
    HIPTRY(uirk_binary(UIRK_ADD, out, s->tempDxt, 0, out, c->nv, c->st), "add");
    HIPTRY(uirk_unary(UIRK_SCALE, out, out, c->nv, s->t->lambda, c->st), "mul_scalar");
    HIPTRY(uirk_binary(UIRK_ADD, out, s->tempAt, 0, out, c->nv, c->st), "add");
    return UIR_OK;
}

/* cgs, :404-585 */
static uir_status sb_cgs(sbctx *s, const float *b, float *x, const float *x0, uint32_t max_iter)
{
    mctx *c = s->c;
    uir_task *t = s->t;
    const size_t nv = c->nv, nb = nv * 4;
    const float tol = 1E-06f;
    const double eps = 2.2204E-16;
    float n2b;
    TRY(sb_l2(s, b, &n2b));
    HIPTRY(hipMemcpyAsync(x, x0, nb, hipMemcpyDeviceToDevice, c->st), "copy");
    int flag = 1;
    float *xmin, *r, *rt, *u, *tmpa, *pp, *ph, *q, *vh, *temp_sum, *uh, *qh;
    TRY(tmp_get(t, 12, nb, &xmin));
    TRY(tmp_get(t, 13, nb, &r));
    TRY(tmp_get(t, 14, nb, &rt));
    TRY(tmp_get(t, 15, nb, &u));
    TRY(tmp_get(t, 16, nb, &tmpa));
    TRY(tmp_get(t, 17, nb, &pp));
    TRY(tmp_get(t, 18, nb, &ph));
    TRY(tmp_get(t, 19, nb, &q));
    TRY(tmp_get(t, 20, nb, &vh));
    TRY(tmp_get(t, 21, nb, &temp_sum));
    TRY(tmp_get(t, 22, nb, &uh));
    TRY(tmp_get(t, 23, nb, &qh));
    HIPTRY(hipMemcpyAsync(xmin, x, nb, hipMemcpyDeviceToDevice, c->st), "copy");
    const float tolb = tol * n2b;
    TRY(sb_process_a(s, x, r));
    HIPTRY(uirk_binary(UIRK_SUB, b, r, 0, r, nv, c->st), "deduction");
    float normr;
    TRY(sb_l2(s, r, &normr));
    float normr_act = normr;
    if (normr <= tolb)
        return UIR_OK; /* :435-440 */
    HIPTRY(hipMemcpyAsync(rt, r, nb, hipMemcpyDeviceToDevice, c->st), "copy");
    float normmin = normr, rho = 1.0f;
    unsigned stag = 0, moresteps = 0;
    const unsigned maxmsteps = 5, maxstagsteps = 3;
    HIPTRY(uirk_set(u, nv, 0, c->st), "set");
    HIPTRY(uirk_set(tmpa, nv, 0, c->st), "set");
    HIPTRY(uirk_set(pp, nv, 0, c->st), "set");
    HIPTRY(uirk_set(q, nv, 0, c->st), "set");
    HIPTRY(uirk_set(temp_sum, nv, 0, c->st), "set");
    for (uint32_t it = 0; it < max_iter; ++it) {
        const float rho1 = rho;
        TRY(sb_dot(s, rt, r, &rho));
        if (rho == 0 || isinf(rho) || isnan(rho)) {
            flag = 4;
            break;
        }
        if (it == 0) {
            HIPTRY(hipMemcpyAsync(u, r, nb, hipMemcpyDeviceToDevice, c->st), "copy");
            HIPTRY(hipMemcpyAsync(pp, u, nb, hipMemcpyDeviceToDevice, c->st), "copy");
        } else {
            const float beta = rho / rho1;
            if (beta == 0 || isinf(beta) || isnan(beta)) {
                flag = 4;
                break;
            }
            HIPTRY(uirk_binary(UIRK_ADD_SCALED, r, q, beta, u, nv, c->st), "add2");
            HIPTRY(uirk_binary(UIRK_ADD_SCALED, q, pp, beta, temp_sum, nv, c->st), "add2");
            HIPTRY(uirk_binary(UIRK_ADD_SCALED, u, temp_sum, beta, pp, nv, c->st), "add2");
        }
        HIPTRY(hipMemcpyAsync(ph, pp, nb, hipMemcpyDeviceToDevice, c->st), "copy");
        TRY(sb_process_a(s, ph, vh));
        float rtvh;
        TRY(sb_dot(s, rt, vh, &rtvh));
        float alpha = 0;
        if (rtvh == 0) {
            flag = 4;
            break;
        }
        alpha = rho / rtvh;
        if (isinf(alpha) || isnan(alpha)) {
            flag = 4;
            break;
        }
        HIPTRY(uirk_binary(UIRK_ADD_SCALED, u, vh, -alpha, q, nv, c->st), "add2");
        HIPTRY(uirk_binary(UIRK_ADD, u, q, 0, uh, nv, c->st), "add");
        float l2uh, l2x;
        TRY(sb_l2(s, uh, &l2uh));
        TRY(sb_l2(s, x, &l2x));
        if (fabs(alpha) * l2uh < eps * l2x) /* :521 */
            stag += 1;
        else
            stag = 0;
        HIPTRY(uirk_binary(UIRK_ADD_SCALED, x, uh, alpha, x, nv, c->st), "add2");
        TRY(sb_process_a(s, uh, qh));
        HIPTRY(uirk_binary(UIRK_ADD_SCALED, r, qh, -alpha, r, nv, c->st), "add2");
        TRY(sb_l2(s, r, &normr));
        normr_act = normr;
        if (normr <= tolb || stag >= maxstagsteps || moresteps) {
            TRY(sb_process_a(s, r, tmpa)); /* :535 applies A to r */
            HIPTRY(uirk_binary(UIRK_SUB, b, tmpa, 0, r, nv, c->st), "deduction");
            TRY(sb_l2(s, r, &normr_act));
            if (normr_act <= tolb) {
                flag = 0;
                break;
            }
            if (stag >= maxstagsteps && moresteps == 0)
                stag = 0;
            moresteps += 1;
            if (moresteps >= maxmsteps) {
                flag = 3;
                break;
            }
        }
        if (normr_act < normmin) {
            normmin = normr_act;
            HIPTRY(hipMemcpyAsync(xmin, x, nb, hipMemcpyDeviceToDevice, c->st), "copy");
        }
        if (stag >= maxstagsteps) {
            flag = 3;
            break;
        }
    }
    if (flag != 0) { /* :566-572 */
        TRY(sb_process_a(s, xmin, tmpa));
        HIPTRY(uirk_binary(UIRK_SUB, b, tmpa, 0, r, nv, c->st), "deduction");
        float l2r;
        TRY(sb_l2(s, r, &l2r));
        if (l2r <= normr_act)
            HIPTRY(hipMemcpyAsync(x, xmin, nb, hipMemcpyDeviceToDevice, c->st), "copy");
    }
    return UIR_OK;
}

/* SBTV process, :230-302 with calculate_b :304-334 and update_db :336-402. */
static uir_status sbtv_run(uir_task *t, float *input, float *u, mctx *c)
{
    const size_t nv = c->nv, ns = c->ns;
    const uint32_t n = c->n;
    sbctx s;
    s.t = t;
    s.c = c;
    float *f, *fbp, *up, *bb, *bx, *by, *dx, *dy, *tmpx, *tmpy, *tmp;
    TRY(tmp_get(t, 0, ns * 4, &s.tempA));
    TRY(tmp_get(t, 1, nv * 4, &s.tempAt));
    TRY(tmp_get(t, 2, nv * 4, &s.tempD));
    TRY(tmp_get(t, 3, nv * 4, &s.tempDxt));
    TRY(tmp_get(t, 4, ns * 4, &f));
    TRY(tmp_get(t, 5, nv * 4, &fbp));
    TRY(tmp_get(t, 6, nv * 4, &up));
    TRY(tmp_get(t, 7, nv * 4, &bb));
    TRY(tmp_get(t, 8, nv * 4, &bx));
    TRY(tmp_get(t, 9, nv * 4, &by));
    TRY(tmp_get(t, 10, nv * 4, &dx));
    TRY(tmp_get(t, 11, nv * 4, &dy));
    /* calculate_b / update_db temporaries (slot 24; CGS owns slots 12-23) */
    float *sbuf;
    TRY(tmp_get(t, 24, nv * 4 * 3, &sbuf));
    tmpx = sbuf;
    tmpy = sbuf + nv;
    tmp = sbuf + 2 * nv;
    uir_workspace *ws = &t->res->ws;
    TRY(uir_ws_normalize(ws, input, ns, c->st));                                  /* :239 */
    HIPTRY(hipMemcpyAsync(f, input, ns * 4, hipMemcpyDeviceToDevice, c->st), "copy"); /* :246-248 */
    TRY(uir_ws_normalize(ws, f, ns, c->st));
    HIPTRY(uirk_set(fbp, nv, 0.0f, c->st), "set");                                /* :251-253 */
    TRY(proj_bp(c->p, n, f, fbp));
    HIPTRY(uirk_set(u, nv, 0.0f, c->st), "set");
    HIPTRY(uirk_set(bx, nv, 0.0f, c->st), "set");
    HIPTRY(uirk_set(by, nv, 0.0f, c->st), "set");
    HIPTRY(uirk_set(dx, nv, 0.0f, c->st), "set");
    HIPTRY(uirk_set(dy, nv, 0.0f, c->st), "set");
    HIPTRY(uirk_unary(UIRK_SCALE, fbp, fbp, nv, t->mu, c->st), "mul_scalar");     /* :279 */
    const float d_lambda = -1 / t->lambda;                                         /* :357 */
    float *ub;
    for (uint32_t it = 0; it < t->num_iterations; ++it) {
        HIPTRY(hipMemcpyAsync(up, u, nv * 4, hipMemcpyDeviceToDevice, c->st), "copy");
        /* calculate_b: tmpx = Dxt(dx - bx), tmpy = Dyt(dy - by), b = fbp + lambda (tmpx + tmpy) */
        HIPTRY(uirk_binary(UIRK_SUB, dx, bx, 0, tmp, nv, c->st), "deduction");
        HIPTRY(uirk_diff(UIRK_DXT, tmp, tmpx, n, n, c->st), "Dxt");
        HIPTRY(uirk_binary(UIRK_SUB, dy, by, 0, tmp, nv, c->st), "deduction");
        HIPTRY(uirk_diff(UIRK_DYT, tmp, tmpy, n, n, c->st), "Dyt");
        HIPTRY(uirk_binary(UIRK_ADD, tmpx, tmpy, 0, bb, nv, c->st), "add");
        HIPTRY(uirk_unary(UIRK_SCALE, bb, bb, nv, t->lambda, c->st), "mul_scalar");
        HIPTRY(uirk_binary(UIRK_ADD, fbp, bb, 0, bb, nv, c->st), "add");
        TRY(sb_cgs(&s, bb, u, up, 30));                                            /* :287 */
        /* update_db */
        ub = s.tempDxt; /* s (the magnitude) */
        HIPTRY(uirk_diff(UIRK_DX, u, tmpx, n, n, c->st), "Dx");
        HIPTRY(uirk_binary(UIRK_ADD, tmpx, bx, 0, tmpx, nv, c->st), "add");
        HIPTRY(uirk_diff(UIRK_DY, u, tmpy, n, n, c->st), "Dy");
        HIPTRY(uirk_binary(UIRK_ADD, tmpy, by, 0, tmpy, nv, c->st), "add");
        HIPTRY(uirk_binary(UIRK_MUL, tmpx, tmpx, 0, tmp, nv, c->st), "mul");
        HIPTRY(hipMemcpyAsync(ub, tmp, nv * 4, hipMemcpyDeviceToDevice, c->st), "copy");
        HIPTRY(uirk_binary(UIRK_MUL, tmpy, tmpy, 0, tmp, nv, c->st), "mul");
        HIPTRY(uirk_binary(UIRK_ADD, ub, tmp, 0, ub, nv, c->st), "add");
        HIPTRY(uirk_unary(UIRK_SQRT, ub, ub, nv, 0, c->st), "sqrt");
        HIPTRY(uirk_set(tmp, nv, d_lambda, c->st), "set");                        /* temp_s_top */
        HIPTRY(uirk_binary(UIRK_ADD, tmp, ub, 0, tmp, nv, c->st), "add");
        HIPTRY(uirk_set(s.tempD, nv, 0.0f, c->st), "set");                        /* Z */
        HIPTRY(uirk_binary(UIRK_MAX, tmp, s.tempD, 0, tmp, nv, c->st), "max");
        HIPTRY(uirk_set(s.tempAt, nv, 1E-12f, c->st), "set");                     /* e12 */
        HIPTRY(uirk_binary(UIRK_MAX, ub, s.tempAt, 0, s.tempAt, nv, c->st), "max");
        HIPTRY(uirk_binary(UIRK_DIV, tmp, s.tempAt, 0, ub, nv, c->st), "div");    /* tresh */
        HIPTRY(uirk_binary(UIRK_MUL, ub, tmpx, 0, dx, nv, c->st), "mul");
        HIPTRY(uirk_binary(UIRK_MUL, ub, tmpy, 0, dy, nv, c->st), "mul");
        HIPTRY(uirk_binary(UIRK_SUB, tmpx, dx, 0, bx, nv, c->st), "deduction");
        HIPTRY(uirk_binary(UIRK_SUB, tmpy, dy, 0, by, nv, c->st), "deduction");
    }
    return UIR_OK;
}

/* ---- process ---- */

/* ufo_ir_state_dependent_task_process, ufo-ir-state-dependent-task.c:194-212:
 * zero the output, then FP or BP over all runs. */
static uir_status projector_process(uir_task *p, const uir_buffer *inputs, uir_buffer *out)
{
    if (!p->res)
        return uir_set_error(UIR_ERR_SETUP, "task is not set up");
    uir_requisition req;
    TRY(proj_requisition(p, &inputs[0], &req));
    TRY(check_buf(out, req.dims[0], req.dims[1], "output"));
    const size_t n_out = (size_t)req.dims[0] * req.dims[1];
    HIPTRY(uirk_set((float *)out->data, n_out, 0.0f, p->res->stream), "set");
    if (p->is_forward)
        return uir_projector_forward(p, &inputs[0], out);
    return uir_projector_backward(p, &inputs[0], out);
}

uir_status uir_task_time_kernels(uir_task *t, const uir_buffer *inputs, uir_buffer *output, uint32_t stride,
                                 float *fp_us, float *bp_us)
{
    if (!t || !inputs || !output || !fp_us || !bp_us)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (t->kind != K_SART)
        return uir_set_error(UIR_ERR_UNSUPPORTED, "kernel timing is for the sart task");
    if (stride == 0)
        stride = 1;
    if (t->res)
        TRY(uir_check_hip(hipSetDevice(t->res->device), "hipSetDevice"));
    mctx c;
    TRY(method_begin(t, inputs, output, &c));
    const float *b = (const float *)inputs[0].data;
    float *x = (float *)output->data;
    sart_prep sp;
    TRY(sart_prepare(t, x, t->relaxation_factor, 0, 1, &c, &sp));
    HIPTRY(hipMemcpyAsync(sp.sino_tmp, b, c.ns * 4, hipMemcpyDeviceToDevice, c.st), "copy");
    hipEvent_t ev[3] = {NULL, NULL, NULL};
    uir_status s = UIR_OK;
    for (int i = 0; i < 3 && s == UIR_OK; ++i)
        s = uir_check_hip(hipEventCreate(&ev[i]), "hipEventCreate");
    double fp_sum = 0, bp_sum = 0;
    uint32_t timed = 0;
    /* one sweep in order, as sart_sweep launches it; every stride-th angle's
     * FP and BP are each bracketed by events */
    for (uint32_t a = 0; a < c.A && s == UIR_OK; ++a) {
        const int tm = a % stride == 0;
        if (tm)
            s = uir_check_hip(hipEventRecord(ev[0], c.st), "hipEventRecord");
        if (s == UIR_OK)
            s = sart_fp_angle(&c, x, sp.sino_tmp, sp.ray_w, sp.axis, sp.starts, a);
        if (s == UIR_OK && tm)
            s = uir_check_hip(hipEventRecord(ev[1], c.st), "hipEventRecord");
        if (s == UIR_OK)
            s = sart_bp_angle(&c, x, sp.sino_tmp, sp.axis, a);
        if (s == UIR_OK && tm) {
            float f = 0, g = 0;
            s = uir_check_hip(hipEventRecord(ev[2], c.st), "hipEventRecord");
            if (s == UIR_OK)
                s = uir_check_hip(hipEventSynchronize(ev[2]), "hipEventSynchronize");
            if (s == UIR_OK)
                s = uir_check_hip(hipEventElapsedTime(&f, ev[0], ev[1]), "hipEventElapsedTime");
            if (s == UIR_OK)
                s = uir_check_hip(hipEventElapsedTime(&g, ev[1], ev[2]), "hipEventElapsedTime");
            fp_sum += f;
            bp_sum += g;
            ++timed;
        }
    }
    for (int i = 0; i < 3; ++i)
        if (ev[i])
            (void)hipEventDestroy(ev[i]);
    if (s != UIR_OK)
        return s;
    *fp_us = timed ? (float)(1e3 * fp_sum / timed) : 0.0f;
    *bp_us = timed ? (float)(1e3 * bp_sum / timed) : 0.0f;
    return UIR_OK;
}

uir_status uir_task_process(uir_task *t, const uir_buffer *inputs, uir_buffer *output)
{
    if (!t || !inputs || !output)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (t->res)
        TRY(uir_check_hip(hipSetDevice(t->res->device), "hipSetDevice"));
    if (t->kind == K_PROJECTOR)
        return projector_process(t, inputs, output);
    mctx c;
    TRY(method_begin(t, inputs, output, &c));
    const float *b = (const float *)inputs[0].data;
    float *x = (float *)output->data;
    switch (t->kind) {
    case K_SIRT:
        return sirt_run(t, b, x, t->relaxation_factor, t->num_iterations, t->positive_constraint, 0, &c);
    case K_SART:
        return sart_run(t, b, x, t->relaxation_factor, t->num_iterations, 0, &c);
    case K_ASDPOCS:
        return asdpocs_run(t, b, x, &c);
    case K_SBTV:
        return sbtv_run(t, (float *)inputs[0].data, x, &c);
    default:
        return uir_set_error(UIR_ERR_UNSUPPORTED, "unknown task kind");
    }
}
