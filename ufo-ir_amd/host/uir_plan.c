/*
 * uir_plan.c -- work-table construction for the table-driven projectors.
 *
 * FP items: for every subset (a direction run, or one angle) and every
 * detector block of UIRK_FP_DETS, angle blocks of up to the tile's angle
 * count.  An angle block is accepted only if the footprint of its rays over
 * any chunk of UIRK_FP_LINES lines fits the LDS staging window.  Ray
 * positions are affine in the detector index and the line index, so the
 * footprint width over a chunk is a convex function of the chunk's position
 * and attains its maximum at the first or the last chunk; both are evaluated
 * exactly here (with a margin for float stepping).  The line model's ray
 * positions lie half a line step further along than joseph's (`shift`, in
 * line steps).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "uir_internal.h"

static int fp_push(uirk_fp_vec *v, uirk_fp_item it)
{
    if (v->n == v->cap) {
        size_t cap = v->cap ? 2 * v->cap : 256;
        uirk_fp_item *p = (uirk_fp_item *)realloc(v->v, cap * sizeof *p);
        if (!p)
            return -1;
        v->v = p;
        v->cap = cap;
    }
    v->v[v->n++] = it;
    return 0;
}

static int bp_push(uirk_bp_vec *v, uirk_bp_item it)
{
    if (v->n == v->cap) {
        size_t cap = v->cap ? 2 * v->cap : 256;
        uirk_bp_item *p = (uirk_bp_item *)realloc(v->v, cap * sizeof *p);
        if (!p)
            return -1;
        v->v = p;
        v->cap = cap;
    }
    v->v[v->n++] = it;
    return 0;
}

/* Footprint width (in staged positions, as the kernel computes it) of the rays
 * of angles [a0, a0 + k) and detectors [d0, d0 + 64) over lines [l0, l0 + C). */
static int window_width(int n, int n_dets, double axis, const float *s, const float *c, int dir, int a0, int k,
                        int d0, int l0, double shift)
{
    const int C = UIRK_FP_LINES;
    const int d1 = d0 + UIRK_FP_DETS - 1 < n_dets - 1 ? d0 + UIRK_FP_DETS - 1 : n_dets - 1;
    const int l1 = l0 + C - 1;
    double lo = 1e300, hi = -1e300;
    for (int a = a0; a < a0 + k; ++a) {
        double det_step, step;
        if (dir == 0) {
            det_step = -1.0 / s[a];
            step = (double)c[a] / (double)s[a];
        } else {
            det_step = 1.0 / c[a];
            step = (double)s[a] / (double)c[a];
        }
        const int ds[2] = {d0, d1};
        const int ls[2] = {l0, l1};
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                const double t = (double)ds[i] + 0.5 - axis;
                const double p = t * det_step - 0.5 * n * step + 0.5 * n + ((double)ls[j] + shift) * step;
                if (p < lo) lo = p;
                if (p > hi) hi = p;
            }
    }
    /* kernel: pmin = floor(lo - 0.5) - 1, W = floor(hi - 0.5) + 3 - pmin;
     * +2 margin for the float rounding of stepped coordinates */
    return (int)(floor(hi - 0.5) - floor(lo - 0.5)) + 4 + 2;
}

/* Whether the footprint of angles [a0, a0 + k) over detector block d0 fits
 * the staging window in every chunk: the width is convex in the chunk's
 * first line, so the first and the last chunk bound it. */
static int block_fits(int n, int n_dets, float axis, const float *s, const float *c, int dir, int a0, int k, int d0,
                      int wmax, double shift)
{
    const int llast = ((n - 1) / UIRK_FP_LINES) * UIRK_FP_LINES;
    return window_width(n, n_dets, axis, s, c, dir, a0, k, d0, 0, shift) <= wmax &&
           window_width(n, n_dets, axis, s, c, dir, a0, k, d0, llast, shift) <= wmax;
}

/* Lines [*lo, *hi) outside which no ray of the block (angles [a0, a0 + k),
 * detectors [d0, d0 + 64)) samples inside the image.  A sample reads a
 * non-border texel iff its position lies in [-0.5, n + 0.5); positions are
 * affine in the detector and line indices, so each angle's set of lines over
 * the detector block is bounded by its two extreme rays.  A margin of 2
 * positions covers the float rounding of stepped coordinates. */
static void line_range(int n, int n_dets, double axis, const float *s, const float *c, int dir, int a0, int k,
                       int d0, int *lo, int *hi, double shift)
{
    const int d1 = d0 + UIRK_FP_DETS - 1 < n_dets - 1 ? d0 + UIRK_FP_DETS - 1 : n_dets - 1;
    const double plo = -0.5 - 2.0, phi = n + 0.5 + 2.0;
    double lmin = 1e300, lmax = -1e300;
    for (int a = a0; a < a0 + k; ++a) {
        double det_step, step;
        if (dir == 0) {
            det_step = -1.0 / s[a];
            step = (double)c[a] / (double)s[a];
        } else {
            det_step = 1.0 / c[a];
            step = (double)s[a] / (double)c[a];
        }
        const int ds[2] = {d0, d1};
        for (int i = 0; i < 2; ++i) {
            const double p0 = ((double)ds[i] + 0.5 - axis) * det_step - 0.5 * n * step + 0.5 * n + shift * step;
            double l_a, l_b;
            if (fabs(step) < 1e-12) {
                if (p0 < plo || p0 > phi)
                    continue;
                l_a = 0;
                l_b = n;
            } else {
                l_a = (plo - p0) / step;
                l_b = (phi - p0) / step;
                if (l_a > l_b) {
                    const double t = l_a;
                    l_a = l_b;
                    l_b = t;
                }
            }
            if (l_a < lmin) lmin = l_a;
            if (l_b > lmax) lmax = l_b;
        }
    }
    /* Rays between the two extremes are covered: for one angle the entry and
     * exit lines are affine in the detector index, so the union over the
     * block lies within [min, max] of the extreme rays. */
    if (!(lmin <= lmax)) {
        *lo = *hi = 0;
        return;
    }
    double a = floor(lmin) - 1, b = ceil(lmax) + 2;
    if (a < 0) a = 0;
    if (b > n) b = n;
    if (a >= b) {
        *lo = *hi = 0;
        return;
    }
    *lo = (int)a;
    *hi = (int)b;
}

static int plan_fp(uirk_fp_vec *out, int slice, int n, int n_dets, float axis, const float *sin_lut,
                   const float *cos_lut, const uir_subset *subs, uint32_t n_subs, int max_angles, int wmax,
                   double shift)
{
    if (max_angles < 1 || max_angles > UIRK_FP_ANGLES_MAX)
        max_angles = UIRK_FP_ANGLES;
    for (uint32_t si = 0; si < n_subs; ++si) {
        const int off = (int)subs[si].offset, cnt = (int)subs[si].n, dir = subs[si].direction;
        /* Angle blocks are cut per detector block: rays far from the axis
         * spread over more positions per angle, so only the outer blocks
         * need fewer angles to fit the window. */
        for (int d0 = 0; d0 < n_dets; d0 += UIRK_FP_DETS) {
            int a = off;
            while (a < off + cnt) {
                /* Shrink in steps of a wave's angle count (a quarter of the
                 * tile's), so that the item's last wave is full: a wave with
                 * no active angle skips the sampling. */
                const int per_wave = max_angles >= 4 ? max_angles / 4 : 1;
                int k = off + cnt - a < max_angles ? off + cnt - a : max_angles;
                while (k > 1 && !block_fits(n, n_dets, axis, sin_lut, cos_lut, dir, a, k, d0, wmax, shift))
                    k = k > per_wave ? ((k - 1) / per_wave) * per_wave : k - 1;
                uirk_fp_item it;
                memset(&it, 0, sizeof it);
                it.slice = slice;
                it.dir = dir;
                it.a0 = a;
                it.na = k;
                it.d0 = d0;
                line_range(n, n_dets, axis, sin_lut, cos_lut, dir, a, k, d0, &it.l_lo, &it.l_hi, shift);
                if (fp_push(out, it))
                    return -1;
                a += k;
            }
        }
    }
    return 0;
}

int uirk_plan_fp(uirk_fp_vec *out, int slice, int n, int n_dets, float axis, const float *sin_lut,
                 const float *cos_lut, const uir_subset *subs, uint32_t n_subs, int max_angles, int wmax)
{
    return plan_fp(out, slice, n, n_dets, axis, sin_lut, cos_lut, subs, n_subs, max_angles, wmax, 0.0);
}

int uirk_plan_fp_line(uirk_fp_vec *out, int slice, int n, int n_dets, float axis, const float *sin_unit,
                      const float *cos_unit, const uir_subset *subs, uint32_t n_subs, int max_angles, int wmax)
{
    return plan_fp(out, slice, n, n_dets, axis, sin_unit, cos_unit, subs, n_subs, max_angles, wmax, 0.5);
}

int uirk_plan_bp(uirk_bp_vec *out, int slice, int n, int tile_rows)
{
    if (tile_rows != 32 && tile_rows != UIRK_BP_TILE)
        tile_rows = UIRK_BP_TILE;
    for (int y0 = 0; y0 < n; y0 += tile_rows)
        for (int x0 = 0; x0 < n; x0 += UIRK_BP_TILE) {
            uirk_bp_item it;
            memset(&it, 0, sizeof it);
            it.slice = slice;
            it.x0 = x0;
            it.y0 = y0;
            if (bp_push(out, it))
                return -1;
        }
    return 0;
}

void uirk_plan_trig(const float *sin_lut, const float *cos_lut, uint32_t n, float *trig)
{
    for (uint32_t i = 0; i < n; ++i) {
        trig[2 * i] = sin_lut[i];
        trig[2 * i + 1] = cos_lut[i];
    }
}

int uirk_plan_num_xcd(void)
{
    static int nx = 0;
    if (nx == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeNumberOfXccs, dev) == hipSuccess && v > 0)
            nx = v > UIRK_MAX_XCD ? UIRK_MAX_XCD : v;
        else
            nx = 1;
    }
    return nx;
}

int uirk_plan_pair_min_items(void)
{
    const char *env = getenv("UIR_PAIR_MIN_ITEMS"); /* override: 0 pairs always */
    if (env && *env)
        return atoi(env);
    static int n = 0;
    if (n == 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            n = 4 * cus;
        else
            n = 1024;
    }
    return n;
}

int uirk_plan_bp_rows32(size_t items)
{
    return items <= (size_t)uirk_plan_pair_min_items();
}

double uirk_fp_item_cost(const uirk_fp_item *it)
{
    return (double)it->na * (double)(it->l_hi - it->l_lo) + 16.0; /* + epilogue */
}

void uirk_plan_xcd(const double *cost, int n, int n_xcd, uirk_xcd_part *out)
{
    memset(out, 0, sizeof *out);
    if (n_xcd < 1)
        n_xcd = 1;
    if (n_xcd > UIRK_MAX_XCD)
        n_xcd = UIRK_MAX_XCD;
    out->n_xcd = n_xcd;
    double total = 0;
    for (int i = 0; i < n; ++i)
        total += cost ? cost[i] : 1.0;
    double acc = 0;
    int x = 1;
    out->start[0] = 0;
    for (int i = 0; i < n && x < n_xcd; ++i) {
        acc += cost ? cost[i] : 1.0;
        while (x < n_xcd && acc >= total * x / n_xcd)
            out->start[x++] = i + 1;
    }
    while (x <= n_xcd)
        out->start[x++] = n;
}

int uirk_plan_grid_xcd(const uirk_xcd_part *part, int n_items)
{
    if (n_items <= 0)
        return 0;
    if (!part || part->n_xcd <= 1)
        return n_items;
    int most = 0;
    for (int x = 0; x < part->n_xcd; ++x) {
        const int c = part->start[x + 1] - part->start[x];
        most = c > most ? c : most;
    }
    return (most > 0 ? most : 1) * part->n_xcd;
}
