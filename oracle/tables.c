/*
 * ORACLE (test infrastructure) — frequency grid, blackbody photon rates and
 * photoionization cross sections.  Each routine restates the reference's
 * arithmetic in the same order so the tables agree bit for bit with the
 * reference sources compiled under oracle/_ref (tests/test_oracle_tables.py).
 */
#include <math.h>
#include "oracle.h"

/* Physical constants, src/constants.h:14-25 and src/prim_chem/chemistry.h. */
#define ORC_HPLANCK (6.626e-27)
#define ORC_KBOLTZ (1.38066e-16)
#define ORC_PI (3.1415926535897932384626433)
#define ORC_NUL (3.28e15)
#define ORC_HI_LIMIT (1.0)
#define ORC_HeI_LIMIT (1.8125)
#define ORC_HeII_LIMIT (4.0)

/* src/source.c:38-53: nu_i = 10^(log_nu_min + (i+1/2) dlog), the step being
 * (max-min)/(float)NGRID_NU evaluated in double. */
void orc_freq_grid(int ngrid, double log_nu_min, double log_nu_max,
                   double *nu, double *dlog_nu)
{
    double dlog = (log_nu_max - log_nu_min) / (float)ngrid;
    for (int i = 0; i < ngrid; i++)
        nu[i] = pow(10.0, log_nu_min + (i + 0.5) * dlog);
    *dlog_nu = dlog;
}

/* src/source.c:10-21: photon-number blackbody normalised at nu_L. */
static double bb_shape(double nu, double temp)
{
    double x = nu * ORC_NUL;
    double at_nu = (x * x) / (exp(ORC_HPLANCK * nu * ORC_NUL / (ORC_KBOLTZ * temp)) - 1.e0);
    double at_limit = (ORC_NUL * ORC_NUL) / (exp(ORC_HPLANCK * ORC_NUL / (ORC_KBOLTZ * temp)) - 1.e0);
    return at_nu / at_limit;
}

/* src/source.c:59-84: successive-refinement midpoint rule (each call n
 * triples the number of points and reuses the previous estimate). */
typedef struct { double s; } midpoint_state;

static double midpoint_step(midpoint_state *st, double a, double b, int n, double temp)
{
    if (n == 1) {
        st->s = (b - a) * bb_shape(0.5 * (a + b), temp);
        return st->s;
    }
    int it = 1;
    for (int j = 1; j < n - 1; j++) it *= 3;
    double tnm = it;
    double del = (b - a) / (3.0 * tnm);
    double ddel = del + del;
    double x = a + 0.5 * del;
    double sum = 0.0;
    for (int j = 1; j <= it; j++) {
        sum += bb_shape(x, temp);
        x += ddel;
        sum += bb_shape(x, temp);
        x += del;
    }
    st->s = (st->s + (b - a) * sum / tnm) / 3.0;
    return st->s;
}

/* src/source.c:86-101: at most 5 refinements, relative tolerance 1e-8. */
static double bin_integral(double temp, double a, double b)
{
    midpoint_state st = {0.0};
    double cur = 0.0, prev = 0.0;
    for (int p = 1; p <= 5; p++) {
        if (p > 1) prev = cur;
        cur = midpoint_step(&st, a, b, p, temp);
        if (p > 1 && fabs(cur - prev) < 1.0e-8 * fabs(prev)) return cur;
    }
    return cur;
}

/* src/source.c:104-138 (type 0, blackbody). */
void orc_photon_rate_bb(int ngrid, double log_nu_min, double log_nu_max,
                        float temperature, double total_rate, double *out)
{
    double dlog = (log_nu_max - log_nu_min) / (float)ngrid;
    double temp = (double)temperature;
    double norm = bin_integral(temp, 1.0, pow(10.0, log_nu_max));
    for (int i = 0; i < ngrid; i++) {
        double lo = log_nu_min + dlog * (double)i;
        double hi = lo + dlog;
        out[i] = bin_integral(temp, pow(10.0, lo), pow(10.0, hi));
        out[i] *= total_rate / norm;
    }
}

/* src/prim_chem/cross_section.c:20-31: hydrogenic HI cross section. */
double orc_csect_HI(double nu)
{
    if (nu < ORC_HI_LIMIT) return 0.0;
    double eps = sqrt(nu - ORC_HI_LIMIT);
    return 6.3e-18 / (nu * nu * nu * nu) * exp(4.0 - 4.0 * atan(eps) / eps)
           / (1.0 - exp(-2.0 * ORC_PI / eps));
}

/* src/prim_chem/cross_section.c:33-40: two-power-law HeI fit. */
double orc_csect_HeI(double nu)
{
    if (nu < ORC_HeI_LIMIT) return 0.0;
    return 7.83e-18 * (1.66 * pow((nu / ORC_HeI_LIMIT), -2.05)
                       - 0.66 * pow((nu / ORC_HeI_LIMIT), -3.05));
}

/* src/prim_chem/cross_section.c:42-53: hydrogenic HeII (Z = 2). */
double orc_csect_HeII(double nu)
{
    if (nu < ORC_HeII_LIMIT) return 0.0;
    double eps = sqrt(nu / ORC_HeII_LIMIT - 1.0);
    return 6.3e-18 * 256.0 / (nu * nu * nu * nu) * exp(4.0 - 4.0 * atan(eps) / eps)
           / (1.0 - exp(-2.0 * ORC_PI / eps)) / 4.0;
}

/* src/cross_section.c:5-22: FP32 table, three species per bin. */
void orc_cross_sections(int ngrid, const double *nu, float *cs)
{
    for (int i = 0; i < ngrid; i++) {
        cs[3 * i + 0] = orc_csect_HI(nu[i]);
        cs[3 * i + 1] = orc_csect_HeI(nu[i]);
        cs[3 * i + 2] = orc_csect_HeII(nu[i]);
    }
}
