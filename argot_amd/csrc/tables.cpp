// Frequency grid, source spectra and cross-section tables (host code of
// libargot_hip.so), with the reference's entry points and arithmetic:
// setup_freq_param / setup_photon_rate (src/source.c:38-138) and
// setup_cross_section with the cross sections of
// src/prim_chem/cross_section.c:20-53.  Compiled without fast-math and
// without FP contraction so the tables match the reference bit for bit.
#include <math.h>

#include "argot_hip.h"

namespace {

constexpr double kHplanck = 6.626e-27;   // constants.h:17
constexpr double kKboltz = 1.38066e-16;  // constants.h:14
constexpr double kPiC = 3.1415926535897932384626433;
constexpr double kNuL = 3.28e15;         // chemistry.h:15

// Photon-number blackbody shape relative to its value at the Lyman limit.
double planck_photons(double nu, double temp) {
  const double x = nu * kNuL;
  const double num = (x * x) / (exp(kHplanck * nu * kNuL / (kKboltz * temp)) - 1.e0);
  const double den = (kNuL * kNuL) / (exp(kHplanck * kNuL / (kKboltz * temp)) - 1.e0);
  return num / den;
}

// Power law in nu (index = photon-number slope), normalised at nu_L.
double power_law(double nu, double index) { return pow(nu, index); }

using Shape = double (*)(double, double);

// Refinement n of the open midpoint rule on [a,b]: n = 1 evaluates the
// midpoint; each further level triples the points and averages with the
// previous estimate `prev`.
double midpoint_level(Shape f, double a, double b, int n, double param, double prev) {
  if (n == 1) return (b - a) * f(0.5 * (a + b), param);
  int it = 1;
  for (int j = 1; j < n - 1; j++) it *= 3;
  const double tnm = it;
  const double del = (b - a) / (3.0 * tnm);
  const double ddel = del + del;
  double x = a + 0.5 * del;
  double sum = 0.0;
  for (int j = 1; j <= it; j++) {
    sum += f(x, param);
    x += ddel;
    sum += f(x, param);
    x += del;
  }
  return (prev + (b - a) * sum / tnm) / 3.0;
}

// Up to five levels, stopping at a relative change below 1e-8.
double integrate(Shape f, double param, double a, double b) {
  double cur = 0.0, prev = 0.0;
  for (int p = 1; p <= 5; p++) {
    if (p > 1) prev = cur;
    cur = midpoint_level(f, a, b, p, param, cur);
    if (p > 1 && fabs(cur - prev) < 1.0e-8 * fabs(prev)) return cur;
  }
  return cur;
}

double sigma_HI(double nu) {
  if (nu < 1.0) return 0.0;
  const double eps = sqrt(nu - 1.0);
  return 6.3e-18 / (nu * nu * nu * nu) * exp(4.0 - 4.0 * atan(eps) / eps) /
         (1.0 - exp(-2.0 * kPiC / eps));
}

double sigma_HeI(double nu) {
  if (nu < 1.8125) return 0.0;
  return 7.83e-18 * (1.66 * pow((nu / 1.8125), -2.05) - 0.66 * pow((nu / 1.8125), -3.05));
}

double sigma_HeII(double nu) {
  if (nu < 4.0) return 0.0;
  const double eps = sqrt(nu / 4.0 - 1.0);
  return 6.3e-18 * 256.0 / (nu * nu * nu * nu) * exp(4.0 - 4.0 * atan(eps) / eps) /
         (1.0 - exp(-2.0 * kPiC / eps)) / 4.0;
}

}  // namespace

extern "C" {

void setup_freq_param(struct freq_param* freq) {
  freq->log_nu_min = 0.0;   // hydrogen molecules off
  freq->log_nu_max = 1.5;
  freq->dlog_nu = (freq->log_nu_max - freq->log_nu_min) / (float)(ARGOT_NGRID_NU);
  for (int i = 0; i < ARGOT_NGRID_NU; i++)
    freq->nu[i] = pow(10.0, freq->log_nu_min + (i + 0.5) * freq->dlog_nu);
}

void setup_photon_rate(struct freq_param* freq, struct radiation_src* src,
                       double total_photon_rate) {
  Shape f = src->type == 0 ? planck_photons : power_law;
  if (src->type != 0 && src->type != 1) return;
  const double norm = integrate(f, (double)src->param, 1.0, pow(10.0, freq->log_nu_max));
  for (int i = 0; i < ARGOT_NGRID_NU; i++) {
    const double lo = freq->log_nu_min + freq->dlog_nu * (double)i;
    const double hi = lo + freq->dlog_nu;
    src->photon_rate[i] = integrate(f, (double)src->param, pow(10.0, lo), pow(10.0, hi));
    src->photon_rate[i] *= total_photon_rate / norm;
  }
}

void setup_cross_section(struct cross_section* csect, struct freq_param* freq) {
  for (int i = 0; i < ARGOT_NGRID_NU; i++) {
    csect[i].csect_HI = sigma_HI(freq->nu[i]);
    csect[i].csect_HeI = sigma_HeI(freq->nu[i]);
    csect[i].csect_HeII = sigma_HeII(freq->nu[i]);
  }
}

void argot_hip_spectral_grid(int ngrid, double log_nu_min, double log_nu_max, float temperature,
                             double* nu, float* cs, double* spec) {
  if (ngrid <= 0 || !(log_nu_max > log_nu_min) || !nu || !cs || !spec) return;
  const double dl = (log_nu_max - log_nu_min) / ngrid;
  double tot = 0.0;
  for (int i = 0; i < ngrid; i++) {
    nu[i] = pow(10.0, log_nu_min + (i + 0.5) * dl);
    cs[3 * i + 0] = (float)sigma_HI(nu[i]);
    cs[3 * i + 1] = (float)sigma_HeI(nu[i]);
    cs[3 * i + 2] = (float)sigma_HeII(nu[i]);
    const double lo = log_nu_min + dl * i;
    spec[i] = integrate(planck_photons, (double)temperature, pow(10.0, lo), pow(10.0, lo + dl));
    tot += spec[i];
  }
  for (int i = 0; i < ngrid; i++) spec[i] /= tot;
}

}  // extern "C"
