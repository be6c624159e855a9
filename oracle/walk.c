/*
 * ORACLE (test infrastructure) — ray segments and the optical-depth walk.
 *
 * orc_ray_segment restates src/calc_ray_segment.c:15-213 for one rank;
 * orc_walk_fp32 restates src/calc_optical_depth.c:65-227 for one segment,
 * keeping every float/double promotion of the C source (TINY is a double
 * literal, sqrt and fabs are the double functions, fabsf the float one).
 * orc_walk_exact is the FP64 oracle: a DDA on exact ray parameters.
 */
#include <math.h>
#include <float.h>
#include "oracle.h"

#define SEG_TINY (1.0e-30)   /* src/calc_ray_segment.c:12 */
#define WALK_TINY (1.0e-31)  /* src/calc_optical_depth.c:15 */

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* mpi_rank with periodic wrap, src/mpi_rank.c:6-25 (one rank: always 0). */

int orc_ray_segment(const orc_grid *g, const float *src, int ix, int iy,
                    int iz, int maxseg, float *start, float *end)
{
    /* :48-50 target centre (double sum rounded to float) */
    float xt = g->xmin + ((float)ix + 0.5) * g->dx;
    float yt = g->ymin + ((float)iy + 0.5) * g->dy;
    float zt = g->zmin + ((float)iz + 0.5) * g->dz;
    /* :53-59 one domain per axis */
    float dxd = (g->xmax - g->xmin) / (float)1;
    float dyd = (g->ymax - g->ymin) / (float)1;
    float dzd = (g->zmax - g->zmin) / (float)1;
    int ixt = (int)((xt - g->xmin) / dxd);
    int iyt = (int)((yt - g->ymin) / dyd);
    int izt = (int)((zt - g->zmin) / dzd);

    /* :61-105 direction cosines, this file's own variant */
    float dx = xt - src[0], dy = yt - src[1], dz = zt - src[2];
    float dl = sqrt(dx * dx + dy * dy + dz * dz + SEG_TINY);
    float cos_t = dz / dl;
    float sin_t = sqrt(1.0 - cos_t * cos_t);
    float dxy = dl * sin_t;
    float cos_p = dx / (dxy + SEG_TINY);
    float sin_p = dy / (dxy + SEG_TINY);
    float xo = cos_p * sin_t;
    if (fabsf(xo) < SEG_TINY) xo = (xo >= 0.0 ? SEG_TINY : -SEG_TINY);
    float yo = sin_p * sin_t;
    if (fabs(yo) < SEG_TINY) yo = (yo >= 0.0 ? SEG_TINY : -SEG_TINY);
    float zo = cos_t;
    if (fabs(zo) < SEG_TINY) zo = (zo >= 0.0 ? SEG_TINY : -SEG_TINY);
    int xs = xo > 0.e0, ys = yo > 0.e0, zs = zo > 0.e0;
    int xn = xs ? 1 : -1, yn = ys ? 1 : -1, zn = zs ? 1 : -1;

    /* :107-126 sources outside the box enter at a face (last test wins) */
    float out = 0.0;
    if (src[0] < g->xmin) out = (g->xmin - src[0]) / xo;
    if (src[0] > g->xmax) out = (g->xmax - src[0]) / xo;
    if (src[1] < g->ymin) out = (g->ymin - src[1]) / yo;
    if (src[1] > g->ymax) out = (g->ymax - src[1]) / yo;
    if (src[2] < g->zmin) out = (g->zmin - src[2]) / zo;
    if (src[2] > g->zmax) out = (g->zmax - src[2]) / zo;
    float sx, sy, sz;
    if (out != 0.0) {
        sx = src[0] + out * xo;
        sy = src[1] + out * yo;
        sz = src[2] + out * zo;
    } else {
        sx = src[0]; sy = src[1]; sz = src[2];
    }
    int ixs = (int)((sx - g->xmin) / dxd);
    int iys = (int)((sy - g->ymin) / dyd);
    int izs = (int)((sz - g->zmin) / dzd);

    /* :137-192 domain walk: split at domain faces */
    int nseg = 0;
    while (ixs != ixt || iys != iyt || izs != izt) {
        float ddx = g->xmin + dxd * (xs + ixs) - sx;
        float ddy = g->ymin + dyd * (ys + iys) - sy;
        float ddz = g->zmin + dzd * (zs + izs) - sz;
        float rx = ddx / xo, ry = ddy / yo, rz = ddz / zo;
        float rmin = fminf(rx, fminf(ry, rz));
        if (rmin == rx) ixs += xn;
        if (rmin == ry) iys += yn;
        if (rmin == rz) izs += zn;
        float ex = sx + rmin * xo, ey = sy + rmin * yo, ez = sz + rmin * zo;
        if (nseg < maxseg) {
            start[3 * nseg + 0] = sx; start[3 * nseg + 1] = sy; start[3 * nseg + 2] = sz;
            end[3 * nseg + 0] = ex; end[3 * nseg + 1] = ey; end[3 * nseg + 2] = ez;
        }
        sx = ex; sy = ey; sz = ez;
        nseg++;
        if (nseg >= maxseg) break;  /* reference asserts < NSEG_PER_RAY */
    }
    if (nseg < maxseg) {
        start[3 * nseg + 0] = sx; start[3 * nseg + 1] = sy; start[3 * nseg + 2] = sz;
        end[3 * nseg + 0] = xt; end[3 * nseg + 1] = yt; end[3 * nseg + 2] = zt;
        nseg++;
    }
    return nseg;
}

/* FP32 number densities of one cell, src/calc_optical_depth.c:150-156:
 * nHe = nH*HELIUM_FACT is formed in double and rounded to float. */
static inline void cell_n(const orc_grid *g, const orc_fields *f, long c,
                          float *nHI, float *nHeI, float *nHeII)
{
    float nH = f->dens[c];
    *nHI = nH * f->fHI[c];
    float nHe = nH * g->helium_fact;
    *nHeI = nHe * f->fHeI[c];
    *nHeII = nHe * f->fHeII[c];
}

typedef struct {
    int ix, iy, iz, ixe, iye, ize;
    float xo, yo, zo;
    int xs, ys, zs, xn, yn, zn;
    float xc, yc, zc;
} ddastate;

/* src/calc_optical_depth.c:76-143: clamped start/end cells, direction
 * cosines with the double promotions of the C source. */
static void dda_setup(const orc_grid *g, const float *s, const float *e, ddastate *d)
{
    d->ixe = (int)((e[0] - g->xmin) / g->dx);
    d->iye = (int)((e[1] - g->ymin) / g->dy);
    d->ize = (int)((e[2] - g->zmin) / g->dz);
    d->ixe = imax(imin(d->ixe, g->nx - 1), 0);
    d->iye = imax(imin(d->iye, g->ny - 1), 0);
    d->ize = imax(imin(d->ize, g->nz - 1), 0);
    d->ix = (int)((s[0] - g->xmin) / g->dx);
    d->iy = (int)((s[1] - g->ymin) / g->dy);
    d->iz = (int)((s[2] - g->zmin) / g->dz);
    d->ix = imax(imin(d->ix, g->nx - 1), 0);
    d->iy = imax(imin(d->iy, g->ny - 1), 0);
    d->iz = imax(imin(d->iz, g->nz - 1), 0);

    float dx = e[0] - s[0], dy = e[1] - s[1], dz = e[2] - s[2];
    float dl = sqrt(dx * dx + dy * dy + dz * dz + WALK_TINY);
    float dxy = sqrt(dx * dx + dy * dy);
    float cos_t = dz / dl;
    float sin_t = dxy / dl;
    float cos_p = dx / (dxy + WALK_TINY);
    float sin_p = dy / (dxy + WALK_TINY);
    float xo = cos_p * sin_t;
    if (fabsf(xo) < WALK_TINY) xo = (xo >= 0.0 ? WALK_TINY : -WALK_TINY);
    float yo = sin_p * sin_t;
    if (fabs(yo) < WALK_TINY) yo = (yo >= 0.0 ? WALK_TINY : -WALK_TINY);
    float zo = cos_t;
    if (fabs(zo) < WALK_TINY) zo = (zo >= 0.0 ? WALK_TINY : -WALK_TINY);
    d->xo = xo; d->yo = yo; d->zo = zo;
    d->xs = xo > 0.e0; d->ys = yo > 0.e0; d->zs = zo > 0.e0;
    d->xn = d->xs ? 1 : -1; d->yn = d->ys ? 1 : -1; d->zn = d->zs ? 1 : -1;
    d->xc = s[0] - g->xmin;
    d->yc = s[1] - g->ymin;
    d->zc = s[2] - g->zmin;
}

static inline int dda_inside(const orc_grid *g, const ddastate *d)
{
    return (d->ix != d->ixe || d->iy != d->iye || d->iz != d->ize) &&
           (0 <= d->ix && d->ix < g->nx) && (0 <= d->iy && d->iy < g->ny) &&
           (0 <= d->iz && d->iz < g->nz);
}

/* One DDA step, src/calc_optical_depth.c:163-179: returns the chord rmin
 * and advances every axis whose exit distance equals it. */
static inline float dda_step(const orc_grid *g, ddastate *d)
{
    float ddx = g->dx * (d->xs + d->ix) - d->xc;
    float ddy = g->dy * (d->ys + d->iy) - d->yc;
    float ddz = g->dz * (d->zs + d->iz) - d->zc;
    float rx = ddx / d->xo, ry = ddy / d->yo, rz = ddz / d->zo;
    float rmin = fminf(rx, fminf(ry, rz));
    if (rmin == rx) d->ix += d->xn;
    if (rmin == ry) d->iy += d->yn;
    if (rmin == rz) d->iz += d->zn;
    d->xc += rmin * d->xo;
    d->yc += rmin * d->yo;
    d->zc += rmin * d->zo;
    return rmin;
}

/* src/calc_optical_depth.c:194-203: remaining length to the end point. */
static inline float dda_depth(const orc_grid *g, const ddastate *d, const float *e)
{
    float x = d->xc + g->xmin, y = d->yc + g->ymin, z = d->zc + g->zmin;
    float ex = e[0] - x, ey = e[1] - y, ez = e[2] - z;
    return sqrt(ex * ex + ey * ey + ez * ez);
}

int orc_walk_fp32(const orc_grid *g, const orc_fields *f, const float *start,
                  const float *end, float *tau, int cap, int *cells)
{
    ddastate d;
    dda_setup(g, start, end, &d);
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
    int nc = 0;
    while (dda_inside(g, &d)) {
        long c = d.iz + (long)g->nz * (d.iy + (long)g->ny * d.ix);
        float a, b, h;
        cell_n(g, f, c, &a, &b, &h);
        if (cells && nc < cap) cells[nc] = (int)c;
        nc++;
        float rmin = dda_step(g, &d);
        t0 += rmin * a;
        t1 += rmin * b;
        t2 += rmin * h;
    }
    float depth = dda_depth(g, &d, end);
    long ce = d.ize + (long)g->nz * (d.iye + (long)g->ny * d.ixe);
    float a, b, h;
    cell_n(g, f, ce, &a, &b, &h);
    if (cells && nc < cap) cells[nc] = (int)ce;
    nc++;
    t0 += depth * a;
    t1 += depth * b;
    t2 += depth * h;
    tau[0] = t0; tau[1] = t1; tau[2] = t2;
    return nc;
}

/* Kahan-compensated variant: the same per-cell terms rmin*n, summed with a
 * running compensation (Kahan 1965). */
static inline void kahan_add(float *sum, float *comp, float term)
{
    float y = term - *comp;
    float t = *sum + y;
    *comp = (t - *sum) - y;
    *sum = t;
}

int orc_walk_fp32_kahan(const orc_grid *g, const orc_fields *f,
                        const float *start, const float *end, float *tau)
{
    ddastate d;
    dda_setup(g, start, end, &d);
    float s[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    int nc = 0;
    while (dda_inside(g, &d)) {
        long cell = d.iz + (long)g->nz * (d.iy + (long)g->ny * d.ix);
        float a, b, h;
        cell_n(g, f, cell, &a, &b, &h);
        nc++;
        float rmin = dda_step(g, &d);
        kahan_add(&s[0], &c[0], rmin * a);
        kahan_add(&s[1], &c[1], rmin * b);
        kahan_add(&s[2], &c[2], rmin * h);
    }
    float depth = dda_depth(g, &d, end);
    long ce = d.ize + (long)g->nz * (d.iye + (long)g->ny * d.ixe);
    float a, b, h;
    cell_n(g, f, ce, &a, &b, &h);
    nc++;
    kahan_add(&s[0], &c[0], depth * a);
    kahan_add(&s[1], &c[1], depth * b);
    kahan_add(&s[2], &c[2], depth * h);
    tau[0] = s[0]; tau[1] = s[1]; tau[2] = s[2];
    return nc;
}

/* FP64 oracle walk: parametrise p(t) = s + t (e - s), t in [0,1]; step
 * through the cell faces with exact double arithmetic on the ray
 * parameter, accumulating n_s * (t_out - t_in) * |e - s|. */
int orc_walk_exact(const orc_grid *g, const orc_fields *f, const float *start,
                   const float *end, double *tau)
{
    double s[3] = {start[0] - (double)g->xmin, start[1] - (double)g->ymin,
                   start[2] - (double)g->zmin};
    double e[3] = {end[0] - (double)g->xmin, end[1] - (double)g->ymin,
                   end[2] - (double)g->zmin};
    double h[3] = {g->dx, g->dy, g->dz};
    int nn[3] = {g->nx, g->ny, g->nz};
    double dv[3] = {e[0] - s[0], e[1] - s[1], e[2] - s[2]};
    double len = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    tau[0] = tau[1] = tau[2] = 0.0;
    int cell[3], step[3];
    double tnext[3], tdelta[3];
    for (int a = 0; a < 3; a++) {
        int c = (int)floor(s[a] / h[a]);
        if (c < 0) c = 0;
        if (c > nn[a] - 1) c = nn[a] - 1;
        cell[a] = c;
        if (dv[a] > 0) {
            step[a] = 1;
            tnext[a] = ((c + 1) * h[a] - s[a]) / dv[a];
            tdelta[a] = h[a] / dv[a];
        } else if (dv[a] < 0) {
            step[a] = -1;
            tnext[a] = (c * h[a] - s[a]) / dv[a];
            tdelta[a] = -h[a] / dv[a];
        } else {
            step[a] = 0;
            tnext[a] = INFINITY;
            tdelta[a] = INFINITY;
        }
    }
    double t = 0.0;
    int ncell = 0;
    for (;;) {
        double tn = tnext[0];
        if (tnext[1] < tn) tn = tnext[1];
        if (tnext[2] < tn) tn = tnext[2];
        double tout = tn < 1.0 ? tn : 1.0;
        if (tout > t) {
            long c = cell[2] + (long)g->nz * (cell[1] + (long)g->ny * cell[0]);
            float a, b, hh;
            cell_n(g, f, c, &a, &b, &hh);
            double l = (tout - t) * len;
            tau[0] += (double)a * l;
            tau[1] += (double)b * l;
            tau[2] += (double)hh * l;
            ncell++;
        }
        if (tn >= 1.0) break;
        t = tn;
        int out = 0;
        for (int a = 0; a < 3; a++) {
            if (tnext[a] == tn) {
                cell[a] += step[a];
                tnext[a] += tdelta[a];
                if (cell[a] < 0 || cell[a] >= nn[a]) out = 1;
            }
        }
        if (out) break;
    }
    return ncell;
}
