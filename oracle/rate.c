/*
 * ORACLE (test infrastructure) — photon-conserving photoionization and
 * photoheating rates of one ray, restating src/calc_photoion_rate.c:52-90
 * (delta_length) and :159-425 (the per-ray body and the merge scaling),
 * with __HELIUM__ on, __HYDROGEN_MOL__ and __COSMOLOGICAL__ off.
 */
#include <math.h>
#include "oracle.h"

#define RATE_TINY (1.0e-31)          /* src/calc_photoion_rate.c:16 */
#define RATE_PI (3.1415926535897932384626433)
#define RATE_HNUL (6.626e-27 * 3.28e15)   /* hplanck*nuL */

/* src/calc_photoion_rate.c:52-90: chord of the target cell along the ray
 * direction, min over axes of |delta/cos|. */
float orc_delta_length(const orc_grid *g, float dx, float dy, float dz)
{
    float dl = sqrt(dx * dx + dy * dy + dz * dz + RATE_TINY);
    float dxy = sqrt(dx * dx + dy * dy + RATE_TINY);
    float cos_t = dz / dl;
    float sin_t = dxy / dl;
    float cos_p = dx / dxy;
    float sin_p = dy / dxy;
    float xo = cos_p * sin_t;
    if (fabsf(xo) < RATE_TINY) xo = (xo >= 0.0 ? RATE_TINY : -RATE_TINY);
    float yo = sin_p * sin_t;
    if (fabsf(yo) < RATE_TINY) yo = (yo >= 0.0 ? RATE_TINY : -RATE_TINY);
    float zo = cos_t;
    if (fabsf(zo) < RATE_TINY) zo = (zo >= 0.0 ? RATE_TINY : -RATE_TINY);
    float rx = fabsf(g->dx / xo);
    float ry = fabsf(g->dy / yo);
    float rz = fabsf(g->dz / zo);
    return fminf(rx, fminf(ry, rz));
}

void orc_ray_rate(const orc_grid *g, const orc_fields *f, const float *src,
                  const double *spec, int ix, int iy, int iz,
                  const double *tau_ray, int mode, double *out)
{
    const double col = g->lunit * g->denstonh;
    double tHI = tau_ray[0] * col;
    double tHeI = tau_ray[1] * col;
    double tHeII = tau_ray[2] * col;

    long c = iz + (long)g->nz * (iy + (long)g->ny * ix);
    /* :182-191 target centre and offset from the ray's source */
    float x = g->dx * ((float)ix + 0.5) + g->xmin;
    float y = g->dy * ((float)iy + 0.5) + g->ymin;
    float z = g->dz * ((float)iz + 0.5) + g->zmin;
    float dx = x - src[0], dy = y - src[1], dz = z - src[2];

    /* :207-246 in-cell chord, half-cell shift, shell volume */
    float delta = orc_delta_length(g, dx, dy, dz);
    float dist = sqrt(dx * dx + dy * dy + dz * dz);
    double nH = f->dens[c] * g->denstonh;
    double nHI = nH * f->fHI[c];
    double dHI = nHI * delta * (g->lunit);
    tHI -= 0.5 * dHI;
    double nHe = nH * g->helium_fact;
    double nHeI = nHe * f->fHeI[c];
    double nHeII = nHe * f->fHeII[c];
    double dHeI = nHeI * delta * (g->lunit);
    double dHeII = nHeII * delta * (g->lunit);
    tHeI -= 0.5 * dHeI;
    tHeII -= 0.5 * dHeII;
    double a = dist + 0.5 * delta, b = dist - 0.5 * delta;
    double lu = g->lunit;
    double vsh = 4.0 * RATE_PI / 3.0 * (a * a * a - b * b * b) * (lu * lu * lu);
    double rHI = 1.0 / (nHI * vsh + RATE_TINY);
    double rHeI = 1.0 / (nHeI * vsh + RATE_TINY);
    double rHeII = 1.0 / (nHeII * vsh + RATE_TINY);

    double gHI = 0.0, hHI = 0.0, gHeI = 0.0, gHeII = 0.0, hHeI = 0.0, hHeII = 0.0;
    /* :285-387 frequency loop */
    for (int k = 0; k < g->ngrid; k++) {
        float nu = g->nu[k];
        double sHI = g->cs[3 * k + 0], sHeI = g->cs[3 * k + 1], sHeII = g->cs[3 * k + 2];
        double tau = tHI * sHI;
        double wHI = dHI * sHI;
        double dtot = wHI;
        tau += tHeI * sHeI;
        tau += tHeII * sHeII;
        double wHeI = dHeI * sHeI;
        double wHeII = dHeII * sHeII;
        dtot += wHeI + wHeII;
        double absorbed;
        if (mode == 0)
            absorbed = spec[k] * exp(-tau) * (1.0e0 - exp(-dtot));
        else
            absorbed = spec[k] * exp(-tau) * (-expm1(-dtot));
        wHI /= (dtot + RATE_TINY);
        gHI += wHI * absorbed * rHI;
        if (nu > 1.0) hHI += wHI * absorbed * (nu - 1.0) * rHI;
        wHeI /= (dtot + RATE_TINY);
        wHeII /= (dtot + RATE_TINY);
        gHeI += wHeI * absorbed * rHeI;
        gHeII += wHeII * absorbed * rHeII;
        if (nu > 1.8125) hHeI += wHeI * absorbed * (nu - 1.8125) * rHeI;
        if (nu > 4.0) hHeII += wHeII * absorbed * (nu - 4.0) * rHeII;
    }
    /* :403-412 merge scaling */
    out[0] = gHI;
    out[1] = hHI * RATE_HNUL;
    out[2] = gHeI;
    out[3] = gHeII;
    out[4] = hHeI * RATE_HNUL;
    out[5] = hHeII * RATE_HNUL;
}

/* Unit-spectrum optically thin kernel F_s(tau_th) used by the surrogate:
 * tau_th[s] is the optical depth of species s at its own threshold bin
 * scale (column times sigma_s(first nonzero bin)); out4 = (sum_nu N sigma_s
 * exp(-tau_nu) for s = HI, HeI, HeII, and sum_nu N sigma_HI (nu-1)
 * exp(-tau_nu)). */
void orc_thin_kernel(int ngrid, const double *nu, const float *cs,
                     const double *spec, const double *tau_th, double *out4)
{
    double sth[3] = {0, 0, 0};
    for (int s = 0; s < 3; s++)
        for (int k = 0; k < ngrid; k++)
            if (cs[3 * k + s] > 0.0f) { sth[s] = cs[3 * k + s]; break; }
    double col[3];
    for (int s = 0; s < 3; s++) col[s] = sth[s] > 0 ? tau_th[s] / sth[s] : 0.0;
    double acc[4] = {0, 0, 0, 0};
    for (int k = 0; k < ngrid; k++) {
        double t = col[0] * cs[3 * k] + col[1] * cs[3 * k + 1] + col[2] * cs[3 * k + 2];
        double w = spec[k] * exp(-t);
        acc[0] += w * cs[3 * k];
        acc[1] += w * cs[3 * k + 1];
        acc[2] += w * cs[3 * k + 2];
        if (nu[k] > 1.0) acc[3] += w * cs[3 * k] * (nu[k] - 1.0);
    }
    for (int i = 0; i < 4; i++) out4[i] = acc[i];
}
