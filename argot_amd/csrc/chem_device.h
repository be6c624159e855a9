// Device-side primordial chemistry of one cell (gfx950): the HI/HII,
// HeI/HeII/HeIII and electron network with photoionization, integrated
// together with radiative heating and cooling, as the reference's GPU
// chemistry does it (src/step_reaction.cu:402-733, src/step_heatcool.cu:19-181,
// src/heatcool_rate.cu:10-74, src/cooling_rate.cu, src/reaction_rate.cu:33-212)
// for its flag set here: helium on, H2 off, __HEATCOOL__ and __CASE_B__ on.
//
// Every float/double promotion of the reference's expressions is written
// out, and the file is compiled without contraction or fast-math, so each
// rounding sits where the C source puts it.  The transcendental functions
// (exp, log, log10, pow) are the device library's double routines; they may
// differ from a host libm in the last bit, which is the only source of
// difference from a host restatement of the same code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "argot_internal.h"

namespace argot {

namespace chem {

constexpr double kKtoEV = 8.61623e-5;    // K_to_eV, chemistry.h:10
constexpr double kXH = 0.755;            // XHYDROGEN, chemistry.h:23
constexpr double kYHe = 0.245;           // YHELIUM, chemistry.h:24
constexpr double kTiny = 1.0e-31;        // TINY, step_reaction.cu:11
constexpr double kDtFactor = 0.1;        // step_reaction.cu:400
constexpr double kSqrtFact = 1.0488088481;   // sqrt(1.1), step_heatcool.cu:14

// One cell's chemical state, struct prim_chem (13 floats).
struct State {
  float wmol, felec, fHI, fHII, fHeI, fHeII, fHeIII;
  float GammaHI, HeatHI, GammaHeI, GammaHeII, HeatHeI, HeatHeII;
};

// The reference's MIN/MAX macros (run_param.h:238-242), NaN behaviour included.
template <typename T>
__device__ __forceinline__ T min_of(T a, T b) { return a > b ? b : a; }
template <typename T>
__device__ __forceinline__ T max_of(T a, T b) { return a > b ? a : b; }

// 1/((1+x_e) X + Y/4), the WMOL macro of chemistry.h:69 (double).
__device__ __forceinline__ double mean_mass_inv(const State& c) {
  return 1.0 / ((1.0 + (double)c.felec) * kXH + 0.25 * kYHe);
}

// ---------------------------------------------------------------- rates ---

// exp(c0 + c1 L + c2 L^2 + ...) with the powers of L formed by repeated
// multiplication and the terms added left to right (the collisional
// ionization fits of reaction_rate.cu; the signs live in the coefficients).
template <int N>
__device__ __forceinline__ double exp_poly(const double (&c)[N], double L) {
  double s = c[0];
  double p = L;
  s = s + c[1] * p;
#pragma unroll
  for (int i = 2; i < N; i++) {
    p = p * L;
    s = s + c[i] * p;
  }
  return exp(s);
}

// a * lam^p / (1 + (lam/b)^q)^r: the Hui & Gnedin (1997) case-B fits.
__device__ __forceinline__ double hui_gnedin(double a, double lam, double p, double b, double q,
                                             double r) {
  return a * pow(lam, p) / pow(1.0 + pow(lam / b, q), r);
}

// k01 (HI + e -> HII + 2e), reaction_rate.cu:33-62.
__device__ inline double k01(double t) {
  constexpr double c[9] = {-32.71396786375,      13.53655609057,     -5.739328757388,
                           1.563154982022,       -0.2877056004391,   0.03482559773736999,
                           -0.00263197617559,    0.0001119543953861, -2.039149852002e-6};
  return exp_poly(c, log(t * kKtoEV));
}

// k02 (HII + e -> HI), case B, reaction_rate.cu:98-103.
__device__ inline double k02(double t) {
  return hui_gnedin(2.753e-14, 2.0 * 157807.0 / t, 1.5, 2.740, 0.407, 2.242);
}

// k03 (HeI + e -> HeII + 2e), reaction_rate.cu:108-136.
__device__ inline double k03(double t) {
  constexpr double c[9] = {-44.09864886561001, 23.91596563469,      -10.75323019821,
                           3.058038757198,     -0.5685118909884001, 0.06795391233790001,
                           -0.005009056101857001, 0.0002067236157507, -3.649161410833e-6};
  const double tev = t * kKtoEV;
  const double L = log(tev);
  return tev > 0.8 ? exp_poly(c, L) : 0.0;
}

// k04 (HeII + e -> HeI), case B, reaction_rate.cu:152-164.
__device__ inline double k04(double t) {
  if (t > 500.3) return 1.26e-14 * pow(2.0 * 285335.0 / t, 0.75);
  const double logt = log10(t);
  return (1.12e-10 - 2.1e-11 * logt) / sqrt(t);
}

// k05 (HeII + e -> HeIII + 2e), reaction_rate.cu:167-192.
__device__ inline double k05(double t) {
  constexpr double c[9] = {-68.71040990212001, 43.93347632635,      -18.48066993568,
                           4.701626486759002,  -0.7692466334492,    0.08113042097303,
                           -0.005324020628287001, 0.0001975705312221, -3.165581065665e-6};
  const double tev = t * kKtoEV;
  const double L = log(tev);
  return tev > 0.8 ? exp_poly(c, L) : 0.0;
}

// k06 (HeIII + e -> HeII), case B, reaction_rate.cu:200-203.
__device__ inline double k06(double t) {
  return hui_gnedin(1.3765e-14, 2.0 * 631515.0 / t, 1.5, 2.740, 0.407, 2.242);
}

// The six rate coefficients at one temperature, each rounded to float as
// the reference stores them (float K_01 ... K_06).
struct KSet {
  float k1, k2, k3, k4, k5, k6;
};

__device__ inline KSet rate_coefficients(float T) {
  const double t = (double)T;
  return {(float)k01(t), (float)k02(t), (float)k03(t), (float)k04(t), (float)k05(t),
          (float)k06(t)};
}

// ------------------------------------------------------------- cooling ---

// Collisional ionization / excitation shape pre/(1 + sqrt(t 1e-5)) *
// exp(-E/(t 1e-5)) shared by cooling_rate.cu:75-114.
__device__ __forceinline__ double collisional(double pre, double sq5, double t5, double E) {
  return pre / (1.0 + sq5) * exp(-E / t5);
}

// heat - cool for one cell [erg cm^-3 s^-1], calc_heatcool_rate_dev
// (heatcool_rate.cu:10-74) with cooling_rate.cu's case-B functions.
__device__ inline double heatcool_rate(const State& c, float zred, double nH, double T,
                                       double helium_fact) {
  const double nHI = nH * (double)c.fHI;
  const double nHII = nH * (double)c.fHII;
  const double nHe = nH * helium_fact;
  const double nHeI = nHe * (double)c.fHeI;
  const double nHeII = nHe * (double)c.fHeII;
  const double nHeIII = nHe * (double)c.fHeIII;
  const double ne = nH * (double)c.felec;

  double heat = (double)c.HeatHI * nHI;
  heat += (double)c.HeatHeI * nHeI + (double)c.HeatHeII * nHeII;

  const double t5 = T * 1.e-5;
  const double sq = sqrt(T);
  const double sq5 = sqrt(t5);
  // recombination (case B), creh2 / crehe2 / crexihe2 / crehe3
  const double creh2 = hui_gnedin(3.435e-30 * T, 2.0 * 157807.0 / T, 1.97, 2.25, 0.376, 3.72);
  const double crehe2 = T < 980.0 ? 1.7e-26 * pow(T, 0.35) : 1.9e-25 * pow(T / 1.0e3, 0.15);
  const double crexihe2 =
      1.24e-13 / pow(T, 1.5) * exp(-4.7 / t5) * (1.0 + 0.3 * exp(-0.94 / t5));
  const double crehe3 = hui_gnedin(3.435e-30 * T, 2.0 * (631515.0 / T), 1.97, 2.25, 0.376, 3.72);
  // collisional ionization / excitation
  const double cioh1 = collisional(1.27e-21 * sq, sq5, t5, 1.578091);
  const double ciohe1 = collisional(9.38e-22 * sq, sq5, t5, 2.853354);
  const double ciohe2 = collisional(4.95e-22 * sq, sq5, t5, 6.31515);
  const double cexh1 = collisional(7.5e-19, sq5, t5, 1.18348);
  const double cexhe1 = collisional(9.1e-27 * pow(T, -0.1687), sq5, t5, 0.13179);
  const double cexhe2 = collisional(5.54e-17 * pow(T, -0.397), sq5, t5, 4.73638);
  // bremsstrahlung with its Gaunt factor, and Compton cooling off the CMB
  const double lg = 5.5 - log10(T);
  const double gaunt = 1.1 + 0.34 * exp(-lg * lg / 3.0);
  const double brems = 1.42e-27 * gaunt * sq;
  const double z1 = 1.0 + (double)zred;
  const double z14 = z1 * z1 * z1 * z1;
  const double compt = 5.406e-36 * z14 * (T - 2.73 * z1);

  double cool = (creh2 * nHII + cioh1 * nHI + cexh1 * nHI) * ne + brems * ne * nHII + compt * ne;
  cool += ((crehe2 + crexihe2) * nHeII + crehe3 * nHeIII) * ne +
          (ciohe1 * nHeI + ciohe2 * nHeII) * ne + (cexhe1 * ne * nHeI + cexhe2 * nHeII) * ne +
          brems * ne * (nHeII + nHeIII);
  return heat - cool;
}

// ------------------------------------------------------ heating/cooling ---

// Implicit energy update over dtime (code units): bracket the root of
// u_new - u_old - Lambda((u_old+u_new)/2) dt and bisect it to 1e-5
// (advance_heatcool_dev, step_heatcool.cu:19-181).  Returns the
// bisection count.
__device__ inline int advance_heatcool(float dens_cell, float* uene, const State& c,
                                       const ChemParams& p, float dtime) {
  const double eunit = (p.lunit * p.lunit) / (p.tunit * p.tunit);
  const double convfact = p.tunit / eunit * (p.lunit * p.lunit * p.lunit) / p.munit;
  const float wmol_min = (float)(4.0 / (5.0 * kXH + 3.0));
  const float wmol_max = (float)(1.0 / (kXH + 0.25 * kYHe));
  // pow(10, LOG_T_MIN + 0.5) = 10 and pow(10, LOG_T_MAX - 0.5) = 1e8 exactly
  const float uene_min = (float)(10.0 / p.uenetok / (double)wmol_max);
  const float uene_max = (float)(1.0e8 / p.uenetok / (double)wmol_min);
  const float anow3i = (float)(1.0 / (double)(p.anow * p.anow * p.anow));
  const double dens = (double)(dens_cell * anow3i);

  double uene_old = (double)*uene;
  double uene_up = uene_old, uene_lo = uene_old;
  if (uene_old < (double)uene_min) {
    uene_up = uene_lo = uene_old = (double)uene_min;
  } else if (uene_old > (double)uene_max) {
    uene_up = uene_lo = uene_old = (double)uene_max;
  }
  const float wmol = (float)mean_mass_inv(c);
  const double nH = dens * p.denstonh;
  auto rate = [&](double T) {
    return heatcool_rate(c, p.znow, nH, T, p.helium_fact) / dens * convfact;
  };
  double T = (double)*uene * p.uenetok * (double)wmol;
  double heatcool = rate(T);
  const double dt = (double)dtime;
  if (heatcool > 0.0) {
    uene_up = min_of(uene_up * kSqrtFact, (double)uene_max);
    uene_lo = max_of(uene_lo / kSqrtFact, (double)uene_min);
    heatcool = rate(uene_up * p.uenetok * (double)wmol);
    while (uene_up - uene_old - heatcool * dt < 0.0) {
      uene_up *= 1.1;
      uene_lo *= 1.1;
      if (uene_up > (double)uene_max) break;
      heatcool = rate(uene_up * p.uenetok * (double)wmol);
    }
  } else {
    uene_up = min_of(uene_up * kSqrtFact, (double)uene_max);
    uene_lo = max_of(uene_lo / kSqrtFact, (double)uene_min);
    heatcool = rate(uene_lo * p.uenetok * (double)wmol);
    while (uene_lo - uene_old - heatcool * dt > 0.0) {
      uene_up /= 1.1;
      uene_lo /= 1.1;
      if (uene_lo < (double)uene_min) break;
      heatcool = rate(uene_lo * p.uenetok * (double)wmol);
    }
  }
  int iter = 0;
  double uene_new;
  do {
    uene_new = 0.5 * (uene_up + uene_lo);
    heatcool = rate(0.5 * (uene_old + uene_new) * p.uenetok * (double)wmol);
    if (uene_new - uene_old - heatcool * dt > 0.0)
      uene_up = uene_new;
    else
      uene_lo = uene_new;
    iter++;
  } while ((uene_up - uene_lo) > 1.0e-5 * uene_new && iter < 32);
  *uene = (float)uene_new;
  return iter;
}

// ---------------------------------------------------------- network step ---

// Number densities of one cell [cm^-3], each rounded to float as the
// reference's float locals hold them.
struct Dens {
  float nH, nHe, nHI, nHII, nHeI, nHeII, nHeIII, ne;
};

__device__ __forceinline__ float expansion_inv3(float anow) {
  return 1.0f / (anow * anow * anow);
}

__device__ __forceinline__ Dens number_densities(float dens, const State& c, const ChemParams& p) {
  Dens n;
  n.nH = (float)((double)dens * p.denstonh * (double)expansion_inv3(p.anow));
  n.nHI = n.nH * c.fHI;
  n.nHII = n.nH * c.fHII;
  n.nHe = (float)((double)n.nH * p.helium_fact);
  n.nHeI = n.nHe * c.fHeI;
  n.nHeII = n.nHe * c.fHeII;
  n.nHeIII = n.nHe * c.fHeIII;
  n.ne = n.nH * c.felec;
  return n;
}

// Temperature of specific energy u (float, as the reference's float T).
__device__ __forceinline__ float temperature(float u, const State& c, const ChemParams& p) {
  return (float)((double)u * p.uenetok * mean_mass_inv(c));
}

// Safe time step [s] from the electron and HI source/sink balance,
// calc_dtime_dev (step_reaction.cu:18-196).  uene is the mesh's current
// specific energy; K holds the coefficients at its temperature.
__device__ inline float safe_dtime(const Dens& n, const State& c, const KSet& K) {
  const float GHI = c.GammaHI;
  const double GHeI = (double)c.GammaHeI, GHeII = (double)c.GammaHeII;
  double elec_p = (double)(K.k1 * n.nHI * n.ne + GHI * n.nHI);
  double elec_m = (double)(K.k2 * n.nHII * n.ne);
  elec_p += (double)(K.k3 * n.nHeI * n.ne + K.k5 * n.nHeII * n.ne) + GHeI * (double)n.nHeI +
            GHeII * (double)n.nHeII;
  elec_m += (double)(K.k4 * n.nHeII * n.ne + K.k6 * n.nHeIII * n.ne);
  const double hi_p = (double)(K.k2 * n.nHII * n.ne);
  const double hi_m = (double)(K.k1 * n.nHI * n.ne + GHI * n.nHI);
  const double eps = (double)1.e-20f;
  const double dt_elec = (double)n.ne / (fabs(elec_p - elec_m) + eps);
  const double dt_hi = (double)n.nHI / (fabs(hi_p - hi_m) + eps);
  double dt = 0.1 * dt_elec + 0.00001 * dt_hi;
  dt = max_of(dt, 1.0);   // shortest step: 1 s
  return (float)dt;
}

// Implicit (backward-Euler) update of one density: (s dt + n)/(1 + a dt)
// with float source s and sink a, double dt, rounded to float.
__device__ __forceinline__ float implicit_update(float n, float s, float a, double dt) {
  return (float)(((double)s * dt + (double)n) / (1.0 + (double)a * dt));
}


// The reaction network and the heating/cooling integrated together over
// dtime (code units), advance_reaction_and_heatcool_dev
// (step_reaction.cu:402-733), split into its set-up (cell_begin) and its
// sub-steps (cell_substep, while cell_more) so that a kernel can interleave
// cells.  c enters as prev_chem (with the cell's photoionization and
// heating rates) and leaves as the updated state; uene enters as prev_uene.
// mesh_uene is the mesh's current specific energy, which the reference reads
// for the first temperature and for every safe-step estimate.
struct CellRun {
  State c;
  float dens, mesh_uene, uene;
  float nH, nHe, nHI, nHII, nHeI, nHeII, nHeIII, ne;
  float GHI;
  double GHeI, GHeII;
  KSet K;
  double dtime_p, dtime_safe, elapsed;
  int nsub, niter;
};

__device__ inline void cell_begin(CellRun& r, const ChemParams& p, float dtime) {
  r.nsub = r.niter = 0;
  r.dtime_p = (double)dtime * p.tunit;
  const Dens n = number_densities(r.dens, r.c, p);
  r.nH = n.nH;
  r.nHe = n.nHe;
  r.nHI = n.nHI;
  r.nHII = n.nHII;
  r.nHeI = n.nHeI;
  r.nHeII = n.nHeII;
  r.nHeIII = n.nHeIII;
  r.ne = n.ne;
  r.GHI = r.c.GammaHI;
  r.GHeI = (double)r.c.GammaHeI;
  r.GHeII = (double)r.c.GammaHeII;
  r.K = rate_coefficients(temperature(r.mesh_uene, r.c, p));
  // calc_dtime reads the same cell state and temperature here: same K.
  r.dtime_safe = kDtFactor * (double)safe_dtime(n, r.c, r.K);
  if (r.dtime_p < r.dtime_safe) r.dtime_safe = r.dtime_p;
  r.elapsed = 0.0;
}

// The reference sub-steps until the elapsed time reaches dtime.  Two
// guards stop a cell earlier, where the reference would run for an
// unbounded time: a state that has turned NaN (e.g. an infinite rate in a
// cell that holds a source) makes every later safe step the 1-second floor,
// and a cell is given at most kMaxSubsteps sub-steps.  Both are counted
// (cell_stopped).
constexpr int kMaxSubsteps = 1000000;

__device__ __forceinline__ bool cell_nan(const CellRun& r) {
  return !(r.c.felec == r.c.felec) || !(r.uene == r.uene);
}

__device__ __forceinline__ bool cell_more(const CellRun& r) {
  return r.elapsed < r.dtime_p && r.nsub < kMaxSubsteps && !cell_nan(r);
}

// Did a guard, not the elapsed time, end this cell?
__device__ __forceinline__ bool cell_stopped(const CellRun& r) {
  return r.elapsed < r.dtime_p;
}

__device__ inline void cell_substep(CellRun& r, const ChemParams& p) {
  State& c = r.c;
  const KSet& K = r.K;
  if (r.elapsed + r.dtime_safe > r.dtime_p) r.dtime_safe = r.dtime_p - r.elapsed;
  r.elapsed += r.dtime_safe;
  r.nsub++;
  const double dts = r.dtime_safe;
  float nHI = r.nHI, nHII = r.nHII, nHeI = r.nHeI, nHeII = r.nHeII, nHeIII = r.nHeIII;
  float ne = r.ne;
  const float GHI = r.GHI;
  const double GHeI = r.GHeI, GHeII = r.GHeII;

  nHI = implicit_update(nHI, K.k2 * nHII * ne, K.k1 * ne + GHI, dts);
  nHII = implicit_update(nHII, K.k1 * nHI * ne + GHI * nHI, K.k2 * ne, dts);
  nHeI = implicit_update(nHeI, K.k4 * nHeII * ne, (float)((double)(K.k3 * ne) + GHeI), dts);
  nHeII = implicit_update(
      nHeII, (float)((double)((K.k3 * nHeI + K.k6 * nHeIII) * ne) + GHeI * (double)nHeI),
      (float)((double)((K.k4 + K.k5) * ne) + GHeII), dts);
  nHeIII = implicit_update(nHeIII, (float)((double)(K.k5 * nHeII * ne) + GHeII * (double)nHeII),
                           K.k6 * ne, dts);
  {
    float s = K.k1 * nHI * ne + GHI * nHI;
    float a = K.k2 * nHII;
    s = (float)((double)s + ((double)(K.k3 * nHeI * ne + K.k5 * nHeII * ne) +
                             GHeI * (double)nHeI + GHeII * (double)nHeII));
    a = a + (K.k4 * nHeII + K.k6 * nHeIII);
    ne = implicit_update(ne, s, a, dts);
  }

  // fractions, renormalised per element; electrons from the ions
  c.fHI = nHI / r.nH;
  c.fHII = nHII / r.nH;
  float cf = 1.0f / (c.fHI + c.fHII);
  c.fHI *= cf;
  c.fHII *= cf;
  const double nHe_t = (double)r.nHe + kTiny;
  c.fHeI = (float)((double)nHeI / nHe_t);
  c.fHeII = (float)((double)nHeII / nHe_t);
  c.fHeIII = (float)((double)nHeIII / nHe_t);
  cf = (float)(1.0 / ((double)(c.fHeI + c.fHeII + c.fHeIII) + kTiny));
  c.fHeI *= cf;
  c.fHeII *= cf;
  c.fHeIII *= cf;
  c.felec = c.fHII;
  c.felec = (float)((double)c.felec +
                    (p.helium_fact * ((double)c.fHeII + 2.0 * (double)c.fHeIII) + kTiny));

  r.dtime_safe = dts / p.tunit;
  r.niter += advance_heatcool(r.dens, &r.uene, c, p, (float)r.dtime_safe);

  // densities, temperature and coefficients for the next sub-step
  const Dens n = number_densities(r.dens, c, p);
  r.nHI = n.nHI;
  r.nHII = n.nHII;
  r.nHeI = n.nHeI;
  r.nHeII = n.nHeII;
  r.nHeIII = n.nHeIII;
  r.ne = n.ne;
  const float T = temperature(r.uene, c, p);
  r.K = rate_coefficients(T);
  // calc_dtime takes the temperature of the mesh's (unchanged) energy;
  // its coefficients equal K only when the two temperatures agree.
  const float Tm = temperature(r.mesh_uene, c, p);
  const KSet Km = Tm == T ? r.K : rate_coefficients(Tm);
  r.dtime_safe = kDtFactor * (double)safe_dtime(n, c, Km);
}

// Relative change of the HI photoionization rate and of the electron
// fraction between two states, diff_chem (chemistry.cu:271-287).
__device__ __forceinline__ float chem_change(const State& a, const State& b) {
  const float dH = (float)((double)fabsf(a.GammaHI - b.GammaHI) / ((double)a.GammaHI + 1.0e-30));
  const float de = (float)((double)fabsf(a.felec - b.felec) / ((double)a.felec + 1.0e-30));
  return max_of(dH, de);
}

// Relative change of the specific energy, diff_uene (chemistry.cu:289-295).
__device__ __forceinline__ float uene_change(float u1, float u2) {
  return (float)((double)fabsf(u1 - u2) / ((double)max_of(u1, u2) + 1.0e-30));
}

}  // namespace chem
}  // namespace argot
