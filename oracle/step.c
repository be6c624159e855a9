/*
 * ORACLE (test infrastructure) — one point-source radiation pass over a
 * range of target cells: calc_ARGOT_part for one rank
 * (src/step_radiation_tree.c:214-309): rays per target from the tree
 * (setup_light_ray_range, src/setup_light_ray_tree.c:397-435), segments
 * (calc_ray_segment), the walk (calc_optical_depth), the per-ray tau sum
 * (accum_optical_depth, src/accum_optical_depth.c:14-168; one rank: the
 * local segment's tau is copied to the ray) and the rates merged into the target
 * (calc_photoion_rate, src/calc_photoion_rate.c:403-425).
 */
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "oracle.h"

#define MAXSEG 8

typedef struct {
    uint64_t cap;
    int *kind, *id;
    float *pos;
} raybuf;

static void rb_grow(raybuf *b, uint64_t need)
{
    if (need <= b->cap) return;
    uint64_t c = b->cap ? b->cap : 256;
    while (c < need) c *= 2;
    b->kind = realloc(b->kind, sizeof(int) * c);
    b->id = realloc(b->id, sizeof(int) * c);
    b->pos = realloc(b->pos, sizeof(float) * 3 * c);
    b->cap = c;
}

static uint64_t rays_for(const orc_tree *t, const orc_grid *g, int ix, int iy, int iz, raybuf *b)
{
    uint64_t n = orc_rays_to(t, g, ix, iy, iz, b->cap, b->kind, b->id, b->pos);
    if (n > b->cap) {
        rb_grow(b, n);
        n = orc_rays_to(t, g, ix, iy, iz, b->cap, b->kind, b->id, b->pos);
    }
    return n;
}

static inline void cell_of(const orc_grid *g, int64_t im, int *ix, int *iy, int *iz)
{
    int64_t nyz = (int64_t)g->ny * g->nz;
    *ix = (int)(im / nyz);
    *iy = (int)((im - *ix * nyz) / g->nz);
    *iz = (int)(im - *ix * nyz - (int64_t)*iy * g->nz);
}

void orc_step(const orc_grid *g, const orc_fields *f, const orc_tree *t,
              int64_t t0, int64_t t1, float *rates_ref, double *rates_f64,
              int kahan, int nthreads, orc_stats *st)
{
    uint64_t nray = 0, ncross = 0, nseg = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : nray, ncross, nseg)
    {
        raybuf b = {0, NULL, NULL, NULL};
        rb_grow(&b, 1024);
#pragma omp for schedule(dynamic, 16)
        for (int64_t im = t0; im < t1; im++) {
            int ix, iy, iz;
            cell_of(g, im, &ix, &iy, &iz);
            uint64_t n = rays_for(t, g, ix, iy, iz, &b);
            float acc[6] = {0, 0, 0, 0, 0, 0};
            double acc64[6] = {0, 0, 0, 0, 0, 0};
            for (uint64_t r = 0; r < n; r++) {
                const float *sp = b.pos + 3 * r;
                const double *spec = orc_ray_spec(t, b.kind[r], b.id[r]);
                float s[3 * MAXSEG], e[3 * MAXSEG];
                int ns = orc_ray_segment(g, sp, ix, iy, iz, MAXSEG, s, e);
                float tau32[3] = {0, 0, 0};
                double tau64[3] = {0, 0, 0};
                for (int q = 0; q < ns; q++) {
                    float ts[3];
                    int nc;
                    if (kahan)
                        nc = orc_walk_fp32_kahan(g, f, s + 3 * q, e + 3 * q, ts);
                    else
                        nc = orc_walk_fp32(g, f, s + 3 * q, e + 3 * q, ts, 0, NULL);
                    /* accum_optical_depth.c:93-101 assigns (not adds) the
                     * local segment's tau to the ray, so with several local
                     * segments the last one processed (serial order) wins */
                    tau32[0] = ts[0]; tau32[1] = ts[1]; tau32[2] = ts[2];
                    ncross += (uint64_t)nc;
                    nseg++;
                    if (rates_f64) {
                        /* the FP64 oracle keeps the same segment: the kept
                         * one is the ray's path inside the box */
                        double te[3];
                        orc_walk_exact(g, f, s + 3 * q, e + 3 * q, te);
                        tau64[0] = te[0]; tau64[1] = te[1]; tau64[2] = te[2];
                    }
                }
                double tr[3] = {tau32[0], tau32[1], tau32[2]};
                double out[6];
                orc_ray_rate(g, f, sp, spec, ix, iy, iz, tr, 0, out);
                for (int k = 0; k < 6; k++) acc[k] = acc[k] + out[k];
                if (rates_f64) {
                    orc_ray_rate(g, f, sp, spec, ix, iy, iz, tau64, 1, out);
                    for (int k = 0; k < 6; k++) acc64[k] += out[k];
                }
            }
            nray += n;
            for (int k = 0; k < 6; k++) rates_ref[6 * im + k] = acc[k];
            if (rates_f64)
                for (int k = 0; k < 6; k++) rates_f64[6 * im + k] = acc64[k];
        }
        free(b.kind); free(b.id); free(b.pos);
    }
    if (st) {
        st->nray = nray;
        st->ncross = ncross;
        st->nseg = nseg;
    }
}

uint64_t orc_ray_records(const orc_grid *g, const orc_fields *f,
                         const orc_tree *t, int64_t t0, int64_t t1,
                         uint64_t cap, int64_t *target, int *kind, int *id,
                         float *tau, int *ncross, uint64_t *hash)
{
    raybuf b = {0, NULL, NULL, NULL};
    rb_grow(&b, 1024);
    int ccap = g->nx + g->ny + g->nz + 8;
    int *cells = malloc(sizeof(int) * ccap);
    uint64_t w = 0;
    for (int64_t im = t0; im < t1; im++) {
        int ix, iy, iz;
        cell_of(g, im, &ix, &iy, &iz);
        uint64_t n = rays_for(t, g, ix, iy, iz, &b);
        for (uint64_t r = 0; r < n; r++) {
            float s[3 * MAXSEG], e[3 * MAXSEG];
            int ns = orc_ray_segment(g, b.pos + 3 * r, ix, iy, iz, MAXSEG, s, e);
            /* the record describes the segment whose tau the ray keeps
             * (the last one; see orc_step) */
            float tt[3];
            int q = ns - 1;
            int nct = orc_walk_fp32(g, f, s + 3 * q, e + 3 * q, tt, ccap, cells);
            uint64_t h = 0xcbf29ce484222325ULL;
            for (int k = 0; k < nct && k < ccap; k++)
                h = (h ^ (uint64_t)(uint32_t)cells[k]) * 0x100000001b3ULL;
            if (w < cap) {
                target[w] = im;
                kind[w] = b.kind[r];
                id[w] = b.id[r];
                tau[3 * w + 0] = tt[0];
                tau[3 * w + 1] = tt[1];
                tau[3 * w + 2] = tt[2];
                ncross[w] = nct;
                hash[w] = h;
            }
            w++;
        }
    }
    free(cells);
    free(b.kind); free(b.id); free(b.pos);
    return w;
}

/* smooth_photoion_rate, src/smooth_photoion_rate.c:1058-1090 (the 27-point
 * stencil of every rate, central weight 30, normalised by 1/56) on one
 * rank: ghost cells copy the nearest interior cell (the source_rank < 0
 * branches, e.g. :180-184, :222-226, :1052-1055).  in/out: six rates per
 * cell in mesh order.  The first thirteen terms add in float; the central
 * term is a double product, so the rest of the sum is double. */
void orc_smooth_rates(int nx, int ny, int nz, const float *in, float *out)
{
    const double normal = 1.0 / (30.0 + 26.0);
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
            for (int iz = 0; iz < nz; iz++) {
                const long c = iz + (long)nz * (iy + (long)ny * ix);
                for (int r = 0; r < 6; r++) {
                    float s = 0.0f;
                    double d = 0.0;
                    int term = 0;
                    for (int dx = -1; dx <= 1; dx++)
                        for (int dy = -1; dy <= 1; dy++)
                            for (int dz = -1; dz <= 1; dz++) {
                                int jx = ix + dx, jy = iy + dy, jz = iz + dz;
                                jx = jx < 0 ? 0 : (jx >= nx ? nx - 1 : jx);
                                jy = jy < 0 ? 0 : (jy >= ny ? ny - 1 : jy);
                                jz = jz < 0 ? 0 : (jz >= nz ? nz - 1 : jz);
                                const float v = in[6 * (jz + (long)nz * (jy + (long)ny * jx)) + r];
                                if (term == 0) s = v;
                                else if (term < 13) s = s + v;
                                else if (term == 13) d = (double)s + 30.0 * (double)v;
                                else d = d + (double)v;
                                term++;
                            }
                    out[6 * c + r] = (float)(normal * d);
                }
            }
}
