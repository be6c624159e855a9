/*
 * ORACLE — test infrastructure, not product code.
 *
 * CPU restatement of ARGOT's ART diffuse-photon transfer on one rank
 * (NNODE = 1), with __HELIUM__ on, __HELIUM_BB__ off and __USE_ATOMIC__ off:
 *   pix2ang_ring          chealpix (Healpix 3.x), restated from the published
 *                         ring-scheme formulas (Gorski et al. 2005, ApJ 622,
 *                         759, sec. 4); the library is not vendored
 *   set_angle_info        src/diffuse_photon/set_angle.c:12-58 (+ :63-179)
 *   sort_angle_id         src/diffuse_photon/set_angle.c:183-247
 *   set_step_func_factor  src/diffuse_photon/set_step_func_factor.c:29-104
 *   calc_rmesh_data       src/diffuse_photon/calc_rmesh_data.c:37-117 and the
 *                         rate fits of src/diffuse_photon/diffuse_chemistry.c
 *   set_position          src/diffuse_photon/set_start_pos.c:195-229, ray
 *                         groups of set_boundary_ray :60-192
 *   calc_diffuse_optical_depth / calc_intensity / ray_cycle
 *                         src/diffuse_photon/ray_tracing.c:60-394
 *   calc_GH_tot, calc_GH_sum
 *                         src/diffuse_photon/calc_GH_tot.c:171-246
 * Built with -ffp-contract=off: every float/double rounding is where the C
 * source puts it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ART_TINY (1.0e-31)
#define ART_PI (3.1415926535897932384626433)
#define ART_HPLANCK (6.626e-27)
#define ART_K_TO_EV (8.61623e-5)
#define ART_EV_TO_ERG (1.6022e-12)
#define ART_NULEV (13.6)
#define ART_XH (0.755)
#define ART_YHE (0.245)

/* cross-section ratios and band integral ratios, diffuse_photon.h:48-75 */
#define RCROSS_HI_nu2 (1.9529766579e-01)
#define RCROSS_HI_nu3 (1.9513732355e-02)
#define RCROSS_HeI_nu3 (2.6858709594e-01)
#define RINT_J_HI_nu2 (1.128855608262e-01)
#define RINT_J_HI_nu3 (5.288790581895e-03)
#define RINT_J_HeI_nu3 (1.2404369898e-01)
#define RINT_H_HI_nu2 (3.1154224364e+00)
#define RINT_H_HI_nu3 (5.2452188546e-01)
#define RINT_H_HeI_nu3 (8.7872474963e+00)

/* ------------------------------------------------------------------ */
/* Directions                                                           */
/* ------------------------------------------------------------------ */

static long isqrt_l(long v)
{
    long r = (long)sqrt((double)v + 0.5);
    while (r * r > v) r--;
    while ((r + 1) * (r + 1) <= v) r++;
    return r;
}

void orc_pix2ang_ring(long nside, long ipix, double *theta, double *phi)
{
    const long nl2 = 2 * nside, nl4 = 4 * nside;
    const long npix = 12 * nside * nside;
    const long ncap = nl2 * (nside - 1);
    const double fact2 = 4.0 / npix;
    double z;
    if (ipix < ncap) {                       /* north polar cap */
        const long iring = (1 + isqrt_l(1 + 2 * ipix)) >> 1;
        const long iphi = ipix + 1 - 2 * iring * (iring - 1);
        z = 1.0 - (double)(iring * iring) * fact2;
        *phi = ((double)iphi - 0.5) * (0.5 * ART_PI) / iring;
    } else if (ipix < npix - ncap) {         /* equatorial belt */
        const long ip = ipix - ncap;
        const long iring = ip / nl4 + nside;
        const long iphi = ip % nl4 + 1;
        const double fodd = ((iring + nside) & 1) ? 1.0 : 0.5;
        z = (double)(nl2 - iring) * (2.0 / (3.0 * nside));
        *phi = ((double)iphi - fodd) * ART_PI / nl2;
    } else {                                 /* south polar cap */
        const long ip = npix - ipix;
        const long iring = (1 + isqrt_l(2 * ip - 1)) >> 1;
        const long iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
        z = -1.0 + (double)(iring * iring) * fact2;
        *phi = ((double)iphi - 0.5) * (0.5 * ART_PI) / iring;
    }
    *theta = acos(z);
}

static int art_base_id(double cos_phi, double sin_phi, double tan_theta, double cos_pi_4)
{
    const double rc = 1.0e0 / fabs(cos_phi), rs = 1.0e0 / fabs(sin_phi);
    if (fabs(cos_phi) >= cos_pi_4) {
        if (0.0e0 <= tan_theta && tan_theta <= rc) return 0;
        if (-rc <= tan_theta && tan_theta < 0.0e0) return 1;
        return cos_phi >= 0.0e0 ? 2 : 3;
    }
    if (0.0e0 <= tan_theta && tan_theta <= rs) return 0;
    if (-rs <= tan_theta && tan_theta < 0.0e0) return 1;
    return sin_phi >= 0.0e0 ? 4 : 5;
}

static int art_corner_id(double cp, double sp, double ct)
{
    if (cp > 0.0e0) {
        if (sp > 0.0e0) return ct > 0.0e0 ? 0 : 1;
        return ct > 0.0e0 ? 2 : 3;
    }
    if (sp > 0.0e0) return ct > 0.0e0 ? 4 : 5;
    return ct > 0.0e0 ? 6 : 7;
}

static void art_swap(float *ovr, short *base, short *corner, int a, int b)
{
    float t[3];
    memcpy(t, ovr + 3 * a, sizeof t);
    memcpy(ovr + 3 * a, ovr + 3 * b, sizeof t);
    memcpy(ovr + 3 * b, t, sizeof t);
    short s = base[a]; base[a] = base[b]; base[b] = s;
    s = corner[a]; corner[a] = corner[b]; corner[b] = s;
}

void orc_art_angles(int nside, int sorted, float *ovr, short *base, short *corner,
                    int *corner_num)
{
    const long npix = 12L * nside * nside;
    const double cos_pi_4 = cos(0.7853981633974483096156608458198757);
    for (int i = 0; i < 8; i++) corner_num[i] = 0;
    for (long ip = 0; ip < npix; ip++) {
        double theta, phi;
        orc_pix2ang_ring(nside, ip, &theta, &phi);
        const double ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi);
        const double tt = tan(theta);
        double x = st * cp, y = st * sp, z = ct;
        if (fabs(x) < ART_TINY) x = (x >= 0.0e0) ? ART_TINY : -ART_TINY;
        if (fabs(y) < ART_TINY) y = (y >= 0.0e0) ? ART_TINY : -ART_TINY;
        if (fabs(z) < ART_TINY) z = (z >= 0.0e0) ? ART_TINY : -ART_TINY;
        ovr[3 * ip + 0] = (float)x;
        ovr[3 * ip + 1] = (float)y;
        ovr[3 * ip + 2] = (float)z;
        base[ip] = (short)art_base_id(cp, sp, tt, cos_pi_4);
        corner[ip] = (short)art_corner_id(cp, sp, ct);
        corner_num[corner[ip]]++;
    }
    if (!sorted) return;
    /* sort_angle_id: two bubble passes, by corner then base within corner */
    for (long i = 0; i < npix - 1; i++)
        for (long j = npix - 1; j > i; j--)
            if (corner[j - 1] > corner[j]) art_swap(ovr, base, corner, (int)j - 1, (int)j);
    for (long i = 0; i < npix - 1; i++)
        for (long j = npix - 1; j > i; j--)
            if (base[j - 1] > base[j] && !(corner[j - 1] < corner[j]))
                art_swap(ovr, base, corner, (int)j - 1, (int)j);
}

/* ------------------------------------------------------------------ */
/* Step-function factors and per-cell absorption / source              */
/* ------------------------------------------------------------------ */

void orc_art_step_factors(int heatcool, float f[6])
{
    const double range = heatcool ? 0.0632353 : 0.0;
    const int nbin = heatcool ? 64 : 1;
    const double unit[3] = {1.0, 1.8125, 4.0};
    const double nuL[3] = {3.2885e15, 5.9483e15, 1.3154e16};
    double (*cs[3])(double) = {orc_csect_HI, orc_csect_HeI, orc_csect_HeII};
    float acc[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    const double block = range / nbin;
    for (int inu = 0; inu < nbin; inu++) {
        for (int s = 0; s < 3; s++) {
            double nu = unit[s] + inu * block;
            double nu_next = unit[s] + (inu + 1) * block;
            if (nu == unit[s]) nu = unit[s] + 1e-7;
            if (nu_next == unit[s]) nu_next = unit[s] + 1e-7;
            double crs[2], crs_next[2];
            crs[0] = 4.0e0 * ART_PI * cs[s](nu) / (ART_HPLANCK * (nu * nuL[s]));
            crs_next[0] = 4.0e0 * ART_PI * cs[s](nu_next) / (ART_HPLANCK * (nu_next * nuL[s]));
            crs[1] = inu == 0 ? 0.0e0 : crs[0] * (nu - unit[s]);
            crs_next[1] = crs[0] * (nu_next - unit[s]);
            acc[s][0] += (crs[0] + crs_next[0]) * 0.5e0;
            acc[s][1] += (crs[1] + crs_next[1]) * 0.5e0;
        }
    }
    for (int s = 0; s < 3; s++) {
        acc[s][0] /= nbin;
        acc[s][1] /= nbin / (ART_HPLANCK * nuL[s]);
        f[2 * s + 0] = acc[s][0];
        f[2 * s + 1] = acc[s][1];
    }
}

/* diffuse_chemistry.c: recombination coefficients */
static double k02_A(double t)
{
    const double tev = t * ART_K_TO_EV;
    const double l1 = log(tev), l2 = l1 * l1, l3 = l2 * l1, l4 = l3 * l1, l5 = l4 * l1,
                 l6 = l5 * l1, l7 = l6 * l1, l8 = l7 * l1, l9 = l8 * l1;
    return exp(-28.61303380689232 - 0.7241125657826851 * l1 - 0.02026044731984691 * l2
               - 0.002380861877349834 * l3 - 0.0003212605213188796 * l4
               - 0.00001421502914054107 * l5 + 4.989108920299513e-6 * l6
               + 5.755614137575758e-7 * l7 - 1.856767039775261e-8 * l8
               - 3.071135243196595e-9 * l9);
}
static double k02_B(double t)
{
    const double lambda = 2.0 * 157807.0 / t;
    return 2.753e-14 * pow(lambda, 1.5) / pow(1.0 + pow(lambda / 2.740, 0.407), 2.242);
}
static double k04_A(double t)
{
    const double tev = t * ART_K_TO_EV;
    if (tev > 0.8)
        return 1.54e-9 * (1. + 0.3 / exp(8.099328789667 / tev))
                   / (exp(40.49664394833662 / tev) * pow(tev, 1.5))
               + 3.92e-13 / pow(tev, 0.6353);
    return 3.92e-13 / pow(tev, 0.6353);
}
static double k04_B(double t)
{
    if (t > 500.3) {
        const double lambda = 2.0 * 285335.0 / t;
        return 1.26e-14 * pow(lambda, 0.75);
    }
    const double logt = log10(t);
    return (1.12e-10 - 2.1e-11 * logt) / sqrt(t);
}
static double k06_A(double t)
{
    return 3.36e-10 / sqrt(t) / pow((t / 1.e3), 0.2) / (1.0 + pow((t / 1.e6), 0.7));
}
static double k06_B(double t)
{
    const double lambda = 2.0 * 631515.0 / t;
    return 1.3765e-14 * pow(lambda, 1.5) / pow(1.0 + pow(lambda / 2.740, 0.407), 2.242);
}

/* The six recombination fits at temperature t (tests): k02_A, k02_B,
 * k04_A, k04_B, k06_A, k06_B. */
void orc_art_recomb(double t, double *out)
{
    out[0] = k02_A(t);
    out[1] = k02_B(t);
    out[2] = k04_A(t);
    out[3] = k04_B(t);
    out[4] = k06_A(t);
    out[5] = k06_B(t);
}

void orc_art_rmesh(long ncell, const float *dens, const float *chem, double denstonh,
                   double uenetok, double helium_fact, float *kap, float *src)
{
    const float csecHI = (float)orc_csect_HI(1.0 + 1.0e-7);
    const float csecHeI = (float)orc_csect_HeI(1.8125 + 1.0e-7);
    const float csecHeII = (float)orc_csect_HeII(4.0 + 1.0e-7);
    for (long c = 0; c < ncell; c++) {
        const float *ch = chem + 7 * c;   /* felec fHI fHII fHeI fHeII fHeIII prev_uene */
        const float wmol = (float)(1.0 / ((1.0 + ch[0]) * ART_XH + 0.25 * ART_YHE));
        float temper = (float)(ch[6] * uenetok * wmol);
        if (temper < 1.0) temper = 1.0;
        const double nH = dens[c] * denstonh;
        const double nHI = ch[1] * nH, nHII = ch[2] * nH;
        double ne = nHII;
        const double nHeI = ch[3] * helium_fact * nH;
        const double nHeII = ch[4] * helium_fact * nH;
        const double nHeIII = ch[5] * helium_fact * nH;
        ne += nHeII + 2.0e0 * nHeIII;
        const double aHI = csecHI * nHI;
        kap[3 * c + 0] = (float)(aHI + ART_TINY);
        const double aHeI = csecHeI * nHeI, aHeII = csecHeII * nHeII;
        kap[3 * c + 1] = (float)(aHI * RCROSS_HI_nu2 + aHeI + ART_TINY);
        kap[3 * c + 2] = (float)(aHI * RCROSS_HI_nu3 + aHeI * RCROSS_HeI_nu3 + aHeII + ART_TINY);
        const double t = temper;
        double em = (k02_A(t) - k02_B(t)) * ne * nHII * ART_NULEV * 1.0 * ART_EV_TO_ERG / (4.0 * ART_PI);
        src[3 * c + 0] = (float)(em / kap[3 * c + 0]);
        em = (k04_A(t) - k04_B(t)) * ne * nHeII * ART_NULEV * 1.8125 * ART_EV_TO_ERG / (4.0 * ART_PI);
        src[3 * c + 1] = (float)(em / kap[3 * c + 1]);
        em = (k06_A(t) - k06_B(t)) * ne * nHeIII * ART_NULEV * 4.0 * ART_EV_TO_ERG / (4.0 * ART_PI);
        src[3 * c + 2] = (float)(em / kap[3 * c + 2]);
    }
}

/* ------------------------------------------------------------------ */
/* Ray tracing of one direction                                         */
/* ------------------------------------------------------------------ */

static void art_intensity(float rmin, long pos, const float *kap, const float *src, float *I,
                          float Iin[3])
{
    for (int b = 0; b < 3; b++) {
        const float tau = kap[3 * pos + b] * rmin;
        const float etau = exp(-tau);
        const float etaum1 = -expm1(-tau);
        I[3 * pos + b] += Iin[b] * etaum1;
        Iin[b] = Iin[b] * etau + src[3 * pos + b] * etaum1;
    }
}

static int art_out(const orc_grid *g, int ix, int iy, int iz)
{
    return ix >= g->nx || iy >= g->ny || iz >= g->nz || ix < 0 || iy < 0 || iz < 0;
}

/* one ray: calc_diffuse_optical_depth's loop body for one start point
 * (one rank: a ray leaving through a side face re-enters at the opposite
 * face with zero intensity) */
static void art_ray(const orc_grid *g, const float ovr[3], int base, float x, float y, float z,
                    const float *kap, const float *src, float *length, float *I, long *ncross)
{
    const float xovr = ovr[0], yovr = ovr[1], zovr = ovr[2];
    const int xs = xovr > 0.0e0f ? 1 : 0, xn = xovr > 0.0e0f ? 1 : -1;
    const int ys = yovr > 0.0e0f ? 1 : 0, yn = yovr > 0.0e0f ? 1 : -1;
    const int zs = zovr > 0.0e0f ? 1 : 0, zn = zovr > 0.0e0f ? 1 : -1;
    float Iin[3] = {0.0f, 0.0f, 0.0f};
    int ix = (int)((x - g->xmin) / g->dx);
    int iy = (int)((y - g->ymin) / g->dy);
    int iz = (int)((z - g->zmin) / g->dz);
    if (xovr < 0.0e0f && ix == g->nx) --ix;
    if (yovr < 0.0e0f && iy == g->ny) --iy;
    if (zovr < 0.0e0f && iz == g->nz) --iz;
    if (art_out(g, ix, iy, iz)) return;
    for (;;) {
        const long pos = iz + (long)g->nz * (iy + (long)g->ny * ix);
        const float dx = g->dx * (xs + ix) - (x - g->xmin);
        const float dy = g->dy * (ys + iy) - (y - g->ymin);
        const float dz = g->dz * (zs + iz) - (z - g->zmin);
        float rdx = dx / xovr, rdy = dy / yovr, rdz = dz / zovr;
        if (rdx < 0.0) rdx = 0.0;
        if (rdy < 0.0) rdy = 0.0;
        if (rdz < 0.0) rdz = 0.0;
        float rmin = fminf(fminf(rdx, rdy), rdz);
        x += rmin * xovr;
        y += rmin * yovr;
        z += rmin * zovr;
        if (rdx == rmin) ix += xn;
        if (rdy == rmin) iy += yn;
        if (rdz == rmin) iz += zn;
        rmin *= g->lunit;
        length[pos] += rmin;
        art_intensity(rmin, pos, kap, src, I, Iin);
        (*ncross)++;
        if (!art_out(g, ix, iy, iz)) continue;
        /* the far face ends the ray; a side face wraps it (one rank) */
        int wrap_x = 0, wrap_y = 0, wrap_z = 0;
        if (base <= 1) {
            if (iz >= g->nz || iz < 0) break;
            wrap_x = 1; wrap_y = 1;
        } else if (base <= 3) {
            if (ix >= g->nx || ix < 0) break;
            wrap_y = 1; wrap_z = 1;
        } else {
            if (iy >= g->ny || iy < 0) break;
            wrap_z = 1; wrap_x = 1;
        }
        /* ray_cycle + reset_I, in the order of the reference's switch */
        for (int step = 0; step < 2; step++) {
            const int axis = base <= 1 ? step : (base <= 3 ? step + 1 : (step == 0 ? 2 : 0));
            if (axis == 0 && wrap_x) {
                if (ix >= g->nx) { ix = 0; x = g->xmin; Iin[0] = Iin[1] = Iin[2] = 0.0f; }
                else if (ix < 0) { ix = g->nx - 1; x = g->xmax; Iin[0] = Iin[1] = Iin[2] = 0.0f; }
            }
            if (axis == 1 && wrap_y) {
                if (iy >= g->ny) { iy = 0; y = g->ymin; Iin[0] = Iin[1] = Iin[2] = 0.0f; }
                else if (iy < 0) { iy = g->ny - 1; y = g->ymax; Iin[0] = Iin[1] = Iin[2] = 0.0f; }
            }
            if (axis == 2 && wrap_z) {
                if (iz >= g->nz) { iz = 0; z = g->zmin; Iin[0] = Iin[1] = Iin[2] = 0.0f; }
                else if (iz < 0) { iz = g->nz - 1; z = g->zmax; Iin[0] = Iin[1] = Iin[2] = 0.0f; }
            }
        }
        if (art_out(g, ix, iy, iz)) break;
    }
}

/* set_position: face-cell start point with the fixed offsets, the
 * starting coordinate on the entry face */
static void art_start(const orc_grid *g, int base, int i, int j, int k, float p[3])
{
    double sx = g->xmin + (double)(i + 0.521212345) * g->dx;
    double sy = g->ymin + (double)(j + 0.508512345) * g->dy;
    double sz = g->zmin + (double)(k + 0.468287655) * g->dz;
    switch (base) {
    case 0: sz = g->zmin; break;
    case 1: sz = g->zmax; break;
    case 2: sx = g->xmin; break;
    case 3: sx = g->xmax; break;
    case 4: sy = g->ymin; break;
    default: sy = g->ymax; break;
    }
    p[0] = (float)sx;
    p[1] = (float)sy;
    p[2] = (float)sz;
}

long orc_art_direction(const orc_grid *g, const float ovr[3], int base, const float *kap,
                       const float *src, float *length, float *I)
{
    const long ncell = (long)g->nx * g->ny * g->nz;
    memset(length, 0, sizeof(float) * ncell);
    memset(I, 0, sizeof(float) * 3 * ncell);
    long ncross = 0;
    /* the four ray groups of set_boundary_ray, traced in group order */
    for (int grp = 0; grp < 4; grp++) {
        const int da = grp >> 1, db = grp & 1;
        if (base <= 1) {
            const int k = base == 0 ? 0 : g->nz - 1;
            for (int i = 0; i < g->nx; i += 2)
                for (int j = 0; j < g->ny; j += 2) {
                    float p[3];
                    art_start(g, base, i + da, j + db, k, p);
                    art_ray(g, ovr, base, p[0], p[1], p[2], kap, src, length, I, &ncross);
                }
        } else if (base <= 3) {
            const int i = base == 2 ? 0 : g->nx - 1;
            for (int j = 0; j < g->ny; j += 2)
                for (int k = 0; k < g->nz; k += 2) {
                    float p[3];
                    art_start(g, base, i, j + da, k + db, p);
                    art_ray(g, ovr, base, p[0], p[1], p[2], kap, src, length, I, &ncross);
                }
        } else {
            const int j = base == 4 ? 0 : g->ny - 1;
            for (int k = 0; k < g->nz; k += 2)
                for (int i = 0; i < g->nx; i += 2) {
                    float p[3];
                    art_start(g, base, i + db, j, k + da, p);
                    art_ray(g, ovr, base, p[0], p[1], p[2], kap, src, length, I, &ncross);
                }
        }
    }
    return ncross;
}

void orc_art_gh_tot(long ncell, const float *length, const float *I, const float *kap,
                    const float *src, const float f[6], float *gh)
{
    for (long c = 0; c < ncell; c++) {
        const float *in = I + 3 * c, *ka = kap + 3 * c, *so = src + 3 * c;
        const double len = length[c];
        double Ig, Ih, sg, sh;
        /* HI */
        Ig = in[0] / ka[0];
        Ig += in[1] * RINT_J_HI_nu2 / ka[1];
        Ig += in[2] * RINT_J_HI_nu3 / ka[2];
        Ih = in[0] / ka[0];
        Ih += in[1] * RINT_H_HI_nu2 / ka[1];
        Ih += in[2] * RINT_H_HI_nu3 / ka[2];
        Ig /= len;
        Ih /= len;
        sg = so[0];
        sg += so[1] * RINT_J_HI_nu2;
        sg += so[2] * RINT_J_HI_nu3;
        sh = so[0];
        sh += so[1] * RINT_H_HI_nu2;
        sh += so[2] * RINT_H_HI_nu3;
        gh[6 * c + 0] += (Ig + sg) * f[0];
        gh[6 * c + 1] += (Ih + sh) * f[1];
        /* HeI */
        Ig = in[1] / ka[1];
        Ig += in[2] * RINT_J_HeI_nu3 / ka[2];
        Ih = in[1] / ka[1];
        Ih += in[2] * RINT_H_HeI_nu3 / ka[2];
        Ig /= len;
        Ih /= len;
        sg = so[1];
        sg += so[2] * RINT_J_HeI_nu3;
        sh = so[1];
        sh += so[2] * RINT_H_HeI_nu3;
        gh[6 * c + 2] += (Ig + sg) * f[2];
        gh[6 * c + 3] += (Ih + sh) * f[3];
        /* HeII */
        Ig = in[2] / ka[2];
        Ih = in[2] / ka[2];
        Ig /= len;
        Ih /= len;
        sg = so[2];
        sh = so[2];
        gh[6 * c + 4] += (Ig + sg) * f[4];
        gh[6 * c + 5] += (Ih + sh) * f[5];
    }
}

long orc_art_run(const orc_grid *g, int nside, int heatcool, const float *dens,
                 const float *chem, double uenetok, float *rates)
{
    const long ncell = (long)g->nx * g->ny * g->nz;
    const long npix = 12L * nside * nside;
    float *ovr = malloc(sizeof(float) * 3 * npix);
    short *base = malloc(sizeof(short) * npix), *corner = malloc(sizeof(short) * npix);
    float *kap = malloc(sizeof(float) * 3 * ncell), *src = malloc(sizeof(float) * 3 * ncell);
    float *length = malloc(sizeof(float) * ncell), *I = malloc(sizeof(float) * 3 * ncell);
    float *gh = calloc(6 * (size_t)ncell, sizeof(float));
    int cnum[8];
    float f[6];
    orc_art_angles(nside, 1, ovr, base, corner, cnum);
    orc_art_step_factors(heatcool, f);
    orc_art_rmesh(ncell, dens, chem, g->denstonh, uenetok, g->helium_fact, kap, src);
    long ncross = 0;
    for (long d = 0; d < npix; d++) {
        ncross += orc_art_direction(g, ovr + 3 * d, base[d], kap, src, length, I);
        orc_art_gh_tot(ncell, length, I, kap, src, f, gh);
    }
    /* calc_GH_sum: mesh order GammaHI, HeatHI, GammaHeI, GammaHeII, HeatHeI, HeatHeII */
    const float r_n_ang = 1.0e0 / npix;
    for (long c = 0; c < ncell; c++) {
        rates[6 * c + 0] += gh[6 * c + 0] * r_n_ang;
        rates[6 * c + 1] += gh[6 * c + 1] * r_n_ang;
        rates[6 * c + 2] += gh[6 * c + 2] * r_n_ang;
        rates[6 * c + 4] += gh[6 * c + 3] * r_n_ang;
        rates[6 * c + 3] += gh[6 * c + 4] * r_n_ang;
        rates[6 * c + 5] += gh[6 * c + 5] * r_n_ang;
    }
    free(ovr); free(base); free(corner); free(kap); free(src); free(length); free(I); free(gh);
    return ncross;
}
