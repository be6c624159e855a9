// ART diffuse-photon transfer on one MI355X (one rank): parallel rays for
// every HEALPix direction sweep the box, carrying three band intensities
// through the cells (emission and absorption per crossing), and the
// per-cell band intensities and path lengths become photoionization and
// heating rates, averaged over directions.
//
// Reference: src/diffuse_photon/ (calc_diffuse_photon_radiation.c:38-120,
// ray_tracing.c:60-394, calc_rmesh_data.c:37-117, calc_GH_tot.c:171-246,
// set_angle.c, set_start_pos.c, set_step_func_factor.c), with __HELIUM__
// on, __HELIUM_BB__ off and __USE_ATOMIC__ off.
//
// GPU layout: directions run in batches; each direction of a batch owns a
// float4 accumulator slice {length, I1, I2, I3} over the mesh, so that the
// rays of all directions in a batch trace in one launch per ray group.
// Within a ray group no two rays touch the same cell (their start points
// are two cells apart and a ray moves at most one cell sideways per layer
// of its dominant axis), so the accumulators need no atomics and the sums
// keep the reference's group order.  Per-cell rates then accumulate over
// the batch's directions in the reference's (sorted) direction order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "argot_device.h"
#include "argot_hip.h"
#include "surrogate.h"

namespace argot {
namespace {

constexpr double kTinyArt = 1.0e-31;
constexpr double kPiArt = 3.1415926535897932384626433;   // constants.h:22
constexpr double kHplanck = 6.626e-27;                   // constants.h:16
constexpr double kKtoeV = 8.61623e-5;                    // chemistry.h:10
constexpr double kEVtoErg = 1.6022e-12;                  // chemistry.h:12
constexpr double kNuLeV = 13.6;                          // chemistry.h:15
constexpr double kXH = 0.755, kYHe = 0.245;              // chemistry.h:23-24

// diffuse_photon.h:48-75 (band cross-section and integral ratios)
constexpr double kRcHI2 = 1.9529766579e-01, kRcHI3 = 1.9513732355e-02;
constexpr double kRcHeI3 = 2.6858709594e-01;
constexpr double kRjHI2 = 1.128855608262e-01, kRjHI3 = 5.288790581895e-03;
constexpr double kRjHeI3 = 1.2404369898e-01;
constexpr double kRhHI2 = 3.1154224364e+00, kRhHI3 = 5.2452188546e-01;
constexpr double kRhHeI3 = 8.7872474963e+00;

// fluid_mesh offsets (156 B): dens @0, prev_chem @96 {wmol, felec, fHI,
// fHII, fHeI, fHeII, fHeIII, GammaHI, HeatHI, GammaHeI, GammaHeII, HeatHeI,
// HeatHeII}, prev_uene @148.
constexpr int kStride = 156;
constexpr int kPrevChem = 96;
constexpr int kPrevUene = 148;

struct ArtDir {
  float ovr[3];
  int base;
  double rcp[3];   // 1 / (double)ovr: exact FP32 quotient dx/ovr via div_by
};

struct ArtGrid {
  int nx, ny, nz;
  int64_t ncell;
  float xmin, ymin, zmin, xmax, ymax, zmax, dx, dy, dz;
  double lunit;
};

// ----------------------------------------------------------------------
// Position of index i along an axis of even length n with the even indices
// first: the rays of one group start two cells apart across the entry face,
// so with this order neighbouring lanes touch neighbouring records.
__device__ __forceinline__ int art_split(int i, int n) { return (i & 1) * (n >> 1) + (i >> 1); }

// Cell index of the layout a direction's slices use: z fastest (split) for
// directions whose dominant axis is x or y; y fastest (split) for z-dominant
// ones (base 0, 1), whose neighbouring rays are two cells apart in y.
__device__ __forceinline__ int64_t art_index(bool zdom, int nx, int ny, int nz, int ix, int iy,
                                             int iz) {
  return zdom ? (int64_t)art_split(iy, ny) + (int64_t)ny * ((int64_t)ix + (int64_t)nx * iz)
              : (int64_t)art_split(iz, nz) + (int64_t)nz * ((int64_t)iy + (int64_t)ny * ix);
}


// Per-cell absorption and source function (calc_rmesh_data.c:37-117).

__device__ inline double k02_A(double t) {
  const double l1 = log(t * kKtoeV), l2 = l1 * l1, l3 = l2 * l1, l4 = l3 * l1, l5 = l4 * l1,
               l6 = l5 * l1, l7 = l6 * l1, l8 = l7 * l1, l9 = l8 * l1;
  return exp(-28.61303380689232 - 0.7241125657826851 * l1 - 0.02026044731984691 * l2 -
             0.002380861877349834 * l3 - 0.0003212605213188796 * l4 -
             0.00001421502914054107 * l5 + 4.989108920299513e-6 * l6 +
             5.755614137575758e-7 * l7 - 1.856767039775261e-8 * l8 -
             3.071135243196595e-9 * l9);
}
__device__ inline double k02_B(double t) {
  const double lambda = 2.0 * 157807.0 / t;
  return 2.753e-14 * pow(lambda, 1.5) / pow(1.0 + pow(lambda / 2.740, 0.407), 2.242);
}
__device__ inline double k04_A(double t) {
  const double tev = t * kKtoeV;
  if (tev > 0.8)
    return 1.54e-9 * (1. + 0.3 / exp(8.099328789667 / tev)) /
               (exp(40.49664394833662 / tev) * pow(tev, 1.5)) +
           3.92e-13 / pow(tev, 0.6353);
  return 3.92e-13 / pow(tev, 0.6353);
}
__device__ inline double k04_B(double t) {
  if (t > 500.3) return 1.26e-14 * pow(2.0 * 285335.0 / t, 0.75);
  return (1.12e-10 - 2.1e-11 * log10(t)) / sqrt(t);
}
__device__ inline double k06_A(double t) {
  return 3.36e-10 / sqrt(t) / pow((t / 1.e3), 0.2) / (1.0 + pow((t / 1.e6), 0.7));
}
__device__ inline double k06_B(double t) {
  const double lambda = 2.0 * 631515.0 / t;
  return 1.3765e-14 * pow(lambda, 1.5) / pow(1.0 + pow(lambda / 2.740, 0.407), 2.242);
}

__global__ void art_rmesh_kernel(const unsigned char* __restrict__ mesh, int nx, int ny, int nz,
                                 double denstonh, double uenetok, double helium_fact,
                                 float csHI, float csHeI, float csHeII,
                                 float4* __restrict__ med, float4* __restrict__ medT) {
  const int64_t ncell = (int64_t)nx * ny * nz;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const unsigned char* m = mesh + c * kStride;
  const float* ch = reinterpret_cast<const float*>(m + kPrevChem);
  const float dens = *reinterpret_cast<const float*>(m);
  const float uene = *reinterpret_cast<const float*>(m + kPrevUene);
  const float wmol = (float)(1.0 / ((1.0 + ch[1]) * kXH + 0.25 * kYHe));
  float temper = (float)((double)uene * uenetok * (double)wmol);
  if (temper < 1.0f) temper = 1.0f;
  const double nH = (double)dens * denstonh;
  const double nHI = (double)ch[2] * nH, nHII = (double)ch[3] * nH;
  double ne = nHII;
  const double nHeI = (double)ch[4] * helium_fact * nH;
  const double nHeII = (double)ch[5] * helium_fact * nH;
  const double nHeIII = (double)ch[6] * helium_fact * nH;
  ne += nHeII + 2.0e0 * nHeIII;
  const double aHI = (double)csHI * nHI;
  const double aHeI = (double)csHeI * nHeI, aHeII = (double)csHeII * nHeII;
  float4 k;
  k.x = (float)(aHI + kTinyArt);
  k.y = (float)(aHI * kRcHI2 + aHeI + kTinyArt);
  k.z = (float)(aHI * kRcHI3 + aHeI * kRcHeI3 + aHeII + kTinyArt);
  k.w = 0.0f;
  const double t = temper;
  float4 s;
  s.x = (float)((k02_A(t) - k02_B(t)) * ne * nHII * kNuLeV * 1.0 * kEVtoErg / (4.0 * kPiArt) /
                (double)k.x);
  s.y = (float)((k04_A(t) - k04_B(t)) * ne * nHeII * kNuLeV * 1.8125 * kEVtoErg /
                (4.0 * kPiArt) / (double)k.y);
  s.z = (float)((k06_A(t) - k06_B(t)) * ne * nHeIII * kNuLeV * 4.0 * kEVtoErg / (4.0 * kPiArt) /
                (double)k.z);
  s.w = 0.0f;
  // the record in both trace layouts
  const int ix = (int)(c / ((int64_t)ny * nz));
  const int iy = (int)((c / nz) % ny);
  const int iz = (int)(c % nz);
  const int64_t cm = art_index(false, nx, ny, nz, ix, iy, iz);
  const int64_t cT = art_index(true, nx, ny, nz, ix, iy, iz);
  med[2 * cm] = k;
  med[2 * cm + 1] = s;
  medT[2 * cT] = k;
  medT[2 * cT + 1] = s;
}


// ----------------------------------------------------------------------
// Ray tracing (ray_tracing.c:60-394 on one rank).

struct TraceArgs {
  ArtGrid g;
  const ArtDir* dirs;     // the batch's directions
  int ndir;
  int group;              // ray group 0..3 (set_boundary_ray's parity classes)
  int nray_max;           // rays per group and direction (largest face / 4)
  const float4* med;      // [2*ncell]: absorption (3 bands), source function (3 bands)
  const float4* medT;     // the same, y-fastest layout (z-dominant directions)
  float4* acc;            // [ndir][ncell] {length, I1, I2, I3}, in the direction's layout
  unsigned long long* ncross;
};

// One band of calc_intensity.  F64: as the reference, exp and expm1 in
// double on the float optical depth, rounded to float.  Otherwise FP32:
// exp(-tau) from the hardware exponential (a few ulp), 1 - exp(-tau) from
// its Taylor series below 0.25 (truncation < 2e-9 relative) so that thin
// cells keep full relative precision.
template <bool F64>
__device__ __forceinline__ void band(float tau, float src, float& Icell, float& Iin) {
  float etau, etaum1;
  if (F64) {
    etau = (float)exp(-(double)tau);
    etaum1 = (float)(-expm1(-(double)tau));
  } else {
    etau = __expf(-tau);
    float p = __builtin_fmaf(tau, -1.0f / 5040, 1.0f / 720);
    p = __builtin_fmaf(tau, p, -1.0f / 120);
    p = __builtin_fmaf(tau, p, 1.0f / 24);
    p = __builtin_fmaf(tau, p, -1.0f / 6);
    p = __builtin_fmaf(tau, p, 1.0f / 2);
    p = __builtin_fmaf(tau, p, -1.0f);
    const float series = -tau * p;   // tau - tau^2/2 + ... + tau^7/5040
    etaum1 = tau < 0.25f ? series : 1.0f - etau;
  }
  Icell += Iin * etaum1;
  Iin = Iin * etau + src * etaum1;
}

template <bool F64>
__global__ __launch_bounds__(256) void art_trace_kernel(TraceArgs a) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int d = (int)(gid / a.nray_max);
  const int r = (int)(gid % a.nray_max);
  unsigned long long nc = 0;
  if (d < a.ndir) {
    const ArtGrid& g = a.g;
    const ArtDir dir = a.dirs[d];
    const int base = dir.base;
    const int da = a.group >> 1, db = a.group & 1;
    int i = 0, j = 0, k = 0, nface = 0;
    if (base <= 1) {
      nface = g.nx * g.ny / 4;
      i = 2 * (r / (g.ny >> 1)) + da;
      j = 2 * (r % (g.ny >> 1)) + db;
      k = base == 0 ? 0 : g.nz - 1;
    } else if (base <= 3) {
      nface = g.ny * g.nz / 4;
      j = 2 * (r / (g.nz >> 1)) + da;
      k = 2 * (r % (g.nz >> 1)) + db;
      i = base == 2 ? 0 : g.nx - 1;
    } else {
      // neighbouring threads take rays two cells apart in z (the fastest
      // axis of the mesh), so that their accesses share cache lines
      nface = g.nz * g.nx / 4;
      i = 2 * (r / (g.nz >> 1)) + db;
      k = 2 * (r % (g.nz >> 1)) + da;
      j = base == 4 ? 0 : g.ny - 1;
    }
    if (r < nface) {
      // set_position (set_start_pos.c:195-229)
      double sx = (double)g.xmin + (double)(i + 0.521212345) * (double)g.dx;
      double sy = (double)g.ymin + (double)(j + 0.508512345) * (double)g.dy;
      double sz = (double)g.zmin + (double)(k + 0.468287655) * (double)g.dz;
      if (base == 0) sz = g.zmin;
      if (base == 1) sz = g.zmax;
      if (base == 2) sx = g.xmin;
      if (base == 3) sx = g.xmax;
      if (base == 4) sy = g.ymin;
      if (base == 5) sy = g.ymax;
      float x = (float)sx, y = (float)sy, z = (float)sz;
      const float xovr = dir.ovr[0], yovr = dir.ovr[1], zovr = dir.ovr[2];
      const int xs = xovr > 0.0f ? 1 : 0, xn = xovr > 0.0f ? 1 : -1;
      const int ys = yovr > 0.0f ? 1 : 0, yn = yovr > 0.0f ? 1 : -1;
      const int zs = zovr > 0.0f ? 1 : 0, zn = zovr > 0.0f ? 1 : -1;
      int ix = (int)((x - g.xmin) / g.dx);
      int iy = (int)((y - g.ymin) / g.dy);
      int iz = (int)((z - g.zmin) / g.dz);
      if (xovr < 0.0f && ix == g.nx) --ix;
      if (yovr < 0.0f && iy == g.ny) --iy;
      if (zovr < 0.0f && iz == g.nz) --iz;
      auto out = [&]() {
        return ix >= g.nx || iy >= g.ny || iz >= g.nz || ix < 0 || iy < 0 || iz < 0;
      };
      float I0 = 0.0f, I1 = 0.0f, I2 = 0.0f;
      const bool zdom = base <= 1;
      float4* __restrict__ acc = a.acc + (int64_t)d * g.ncell;
      const float4* __restrict__ med = zdom ? a.medT : a.med;
      auto cell = [&]() { return art_index(zdom, g.nx, g.ny, g.nz, ix, iy, iz); };
      bool live = !out();
      int64_t pos = 0;
      float4 kp{}, sp{}, ac{};
      if (live) {
        pos = cell();
        kp = med[2 * pos];
        sp = med[2 * pos + 1];
        ac = acc[pos];
      }
      // Each pass: the crossing's geometry and the next cell (with the
      // side-face wrap) first, the next cell's loads issued, then the
      // current cell's intensities (the geometry never reads the medium).
      while (live) {
        const float ddx = g.dx * (float)(xs + ix) - (x - g.xmin);
        const float ddy = g.dy * (float)(ys + iy) - (y - g.ymin);
        const float ddz = g.dz * (float)(zs + iz) - (z - g.zmin);
        float rdx = div_by(ddx, dir.rcp[0]);
        float rdy = div_by(ddy, dir.rcp[1]);
        float rdz = div_by(ddz, dir.rcp[2]);
        if (rdx < 0.0f) rdx = 0.0f;
        if (rdy < 0.0f) rdy = 0.0f;
        if (rdz < 0.0f) rdz = 0.0f;
        float rmin = fminf(fminf(rdx, rdy), rdz);
        x += rmin * xovr;
        y += rmin * yovr;
        z += rmin * zovr;
        if (rdx == rmin) ix += xn;
        if (rdy == rmin) iy += yn;
        if (rdz == rmin) iz += zn;
        rmin = (float)((double)rmin * g.lunit);
        // next cell: far face ends the ray; a side face wraps it to the
        // opposite face, whose intensity restarts at zero after this
        // cell's update (ray_cycle + reset_I on one rank)
        bool wrapped = false, cont = true;
        if (out()) {
          const bool far = base <= 1 ? (iz >= g.nz || iz < 0)
                                     : (base <= 3 ? (ix >= g.nx || ix < 0) : (iy >= g.ny || iy < 0));
          if (far) {
            cont = false;
          } else {
            auto wrap_x = [&]() {
              if (ix >= g.nx) { ix = 0; x = g.xmin; wrapped = true; }
              else if (ix < 0) { ix = g.nx - 1; x = g.xmax; wrapped = true; }
            };
            auto wrap_y = [&]() {
              if (iy >= g.ny) { iy = 0; y = g.ymin; wrapped = true; }
              else if (iy < 0) { iy = g.ny - 1; y = g.ymax; wrapped = true; }
            };
            auto wrap_z = [&]() {
              if (iz >= g.nz) { iz = 0; z = g.zmin; wrapped = true; }
              else if (iz < 0) { iz = g.nz - 1; z = g.zmax; wrapped = true; }
            };
            if (base <= 1) { wrap_x(); wrap_y(); }
            else if (base <= 3) { wrap_y(); wrap_z(); }
            else { wrap_z(); wrap_x(); }
            cont = !out();
          }
        }
        int64_t npos = pos;
        float4 nkp = kp, nsp = sp, nac = ac;
        if (cont) {
          // another ray of this group never touches these cells, and this
          // ray never returns to one: the early read of acc is safe
          npos = cell();
          nkp = med[2 * npos];
          nsp = med[2 * npos + 1];
          nac = acc[npos];
        }
        ac.x += rmin;
        band<F64>(kp.x * rmin, sp.x, ac.y, I0);
        band<F64>(kp.y * rmin, sp.y, ac.z, I1);
        band<F64>(kp.z * rmin, sp.z, ac.w, I2);
        acc[pos] = ac;
        nc++;
        if (wrapped) I0 = I1 = I2 = 0.0f;
        pos = npos;
        kp = nkp;
        sp = nsp;
        ac = nac;
        live = cont;
      }
    }
  }
  // crossings of the wavefront, one atomic per wavefront
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) nc += __shfl_xor(nc, off, 64);
  if ((threadIdx.x & 63) == 0 && nc) atomicAdd(a.ncross, nc);
}

__global__ void art_fill_kernel(float4* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ----------------------------------------------------------------------
// calc_GH_tot for the batch's directions in order: gh[k][c] (FP32, 6
// outputs GHI, HHI, GHeI, HHeI, GHeII, HHeII); optionally Kahan-compensated
// across directions.
// One block per 32 x 32 (y, z) tile of one x plane, one thread per cell.
// The slices of x- and y-dominant directions are read in their own layout
// (z fastest); those of z-dominant directions (y fastest) go through an LDS
// tile, read along y and used along z, so that every read is a contiguous
// run.  Each cell's sums take the directions in batch order.
// Grid: (ceil(ny/32) ceil(nz/32), nx), blocks of kGhTile^2 threads.
constexpr int kGhTile = 32;
template <bool KAHAN>
__global__ __launch_bounds__(kGhTile * kGhTile) void art_gh_kernel(
    const float4* __restrict__ acc, const ArtDir* __restrict__ dirs, int ndir, int nx, int ny,
    int nz, const float4* __restrict__ med, float f0, float f1, float f2, float f3, float f4,
    float f5, float* __restrict__ gh, float* __restrict__ comp) {
  __shared__ float4 tile[kGhTile][kGhTile + 1];
  const int64_t ncell = (int64_t)nx * ny * nz;
  const int ntz = (nz + kGhTile - 1) / kGhTile;
  const int y0 = (int)(blockIdx.x / ntz) * kGhTile, z0 = (int)(blockIdx.x % ntz) * kGhTile;
  const int ix = blockIdx.y;
  const int tz = threadIdx.x % kGhTile, ty = threadIdx.x / kGhTile;
  const int iy = y0 + ty, iz = z0 + tz;
  const bool valid = iy < ny && iz < nz;
  // this cell's record in the layout of x- and y-dominant directions
  const int64_t cs = valid ? art_index(false, nx, ny, nz, ix, iy, iz) : 0;
  const int64_t c = ((int64_t)ix * ny + iy) * nz + iz;
  // the record this thread loads into the tile for z-dominant directions:
  // cell (y0 + tz, z0 + ty), read along y
  const bool tvalid = y0 + tz < ny && z0 + ty < nz;
  const int64_t ct = tvalid ? art_index(true, nx, ny, nz, ix, y0 + tz, z0 + ty) : 0;
  const float4 ka = valid ? med[2 * cs] : make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  const float4 so = valid ? med[2 * cs + 1] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float s[6], e[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    s[k] = valid ? gh[k * ncell + c] : 0.0f;
    e[k] = KAHAN && valid ? comp[k * ncell + c] : 0.0f;
  }
  const double f[6] = {f0, f1, f2, f3, f4, f5};
  // Per-cell reciprocals: the float quotients I/kappa are exact through
  // div_by; the double quotients become products (within an ulp of double).
  const double rkx = 1.0 / (double)ka.x, rky = 1.0 / (double)ka.y, rkz = 1.0 / (double)ka.z;
  // source-function terms do not depend on the direction
  double sg0 = so.x;
  sg0 += (double)so.y * kRjHI2;
  sg0 += (double)so.z * kRjHI3;
  double sh0 = so.x;
  sh0 += (double)so.y * kRhHI2;
  sh0 += (double)so.z * kRhHI3;
  double sg1 = so.y;
  sg1 += (double)so.z * kRjHeI3;
  double sh1 = so.y;
  sh1 += (double)so.z * kRhHeI3;
  const double sg2 = so.z, sh2 = so.z;
  for (int d = 0; d < ndir; d++) {
    float4 a;
    if (dirs[d].base <= 1) {   // the same for the whole block
      __syncthreads();         // the previous tile has been read
      if (tvalid) tile[ty][tz] = acc[(int64_t)d * ncell + ct];
      __syncthreads();
      a = tile[tz][ty];
    } else {
      a = acc[(int64_t)d * ncell + cs];
    }
    if (!valid) continue;
    const double rl = 1.0 / (double)a.x;
    const double q1 = (double)div_by(a.y, rkx);   // I_1 / kappa_1 (float, as the reference)
    const double q2 = (double)div_by(a.z, rky);
    const double q3 = (double)div_by(a.w, rkz);
    const double z2 = (double)a.z * rky, w3 = (double)a.w * rkz;
    double v[6];
    // HI
    double Ig = q1;
    Ig += z2 * kRjHI2;
    Ig += w3 * kRjHI3;
    double Ih = q1;
    Ih += z2 * kRhHI2;
    Ih += w3 * kRhHI3;
    v[0] = (Ig * rl + sg0) * f[0];
    v[1] = (Ih * rl + sh0) * f[1];
    // HeI
    Ig = q2;
    Ig += w3 * kRjHeI3;
    Ih = q2;
    Ih += w3 * kRhHeI3;
    v[2] = (Ig * rl + sg1) * f[2];
    v[3] = (Ih * rl + sh1) * f[3];
    // HeII
    v[4] = (q3 * rl + sg2) * f[4];
    v[5] = (q3 * rl + sh2) * f[5];
#pragma unroll
    for (int k = 0; k < 6; k++) {
      if (KAHAN) {
        const float y = (float)v[k] - e[k];
        const float t = s[k] + y;
        e[k] = (t - s[k]) - y;
        s[k] = t;
      } else {
        s[k] = (float)((double)s[k] + v[k]);
      }
    }
  }
  if (!valid) return;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    gh[k * ncell + c] = s[k];
    if (KAHAN) comp[k * ncell + c] = e[k];
  }
}

// calc_GH_sum: prev_chem rates += gh * (float)(1/N_ANG)
__global__ void art_gh_sum_kernel(unsigned char* __restrict__ mesh, int64_t ncell,
                                  const float* __restrict__ gh, float r_n_ang) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  float* r = reinterpret_cast<float*>(mesh + c * kStride + kPrevChem + 28);
  r[0] += gh[0 * ncell + c] * r_n_ang;   // GammaHI  += GHI_tot
  r[1] += gh[1 * ncell + c] * r_n_ang;   // HeatHI   += HHI_tot
  r[2] += gh[2 * ncell + c] * r_n_ang;   // GammaHeI += GHeI_tot
  r[4] += gh[3 * ncell + c] * r_n_ang;   // HeatHeI  += HHeI_tot
  r[3] += gh[4 * ncell + c] * r_n_ang;   // GammaHeII += GHeII_tot
  r[5] += gh[5 * ncell + c] * r_n_ang;   // HeatHeII += HHeII_tot
}

// ----------------------------------------------------------------------
// Host side.

int g_nside = 8;             // N_SIDE (run_param.h:159)
int g_heatcool = 1;          // __HEATCOOL__ (src/Makefile:187)
int g_batch = 0;             // directions per batch; 0: as many as 8 GiB of slices hold, <= 128
int g_kahan = 0;
int g_exp64 = 0;             // 1: exp/expm1 in FP64 as the reference
argot_diffuse_stats g_dstats{};

struct ArtCtx {
  int nside = 0;
  std::vector<ArtDir> dirs;       // sorted order
  float fact[6] = {0, 0, 0, 0, 0, 0};
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  ArtDir* d_dirs = nullptr;
  float4 *med = nullptr, *acc = nullptr;   // med: [2*ncell] absorption, source
  float4* medT = nullptr;   // y-fastest medium
  float *gh = nullptr, *comp = nullptr;
  unsigned long long* ncross = nullptr;
  int64_t ncell = 0;
  int batch = 0;
};
ArtCtx a_;

#define A_OK(expr)                                                                  \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      argot_hip_internal_set_error(_e == hipErrorOutOfMemory ? ARGOT_ERR_NOMEM      \
                                                             : ARGOT_ERR_HIP);      \
      return;                                                                       \
    }                                                                               \
  } while (0)

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// HEALPix ring scheme (Gorski et al. 2005, ApJ 622, 759, sec. 4): centre
// of pixel ipix as (theta, phi), the role of chealpix's pix2ang_ring.
long isqrt_l(long v) {
  long r = (long)sqrt((double)v + 0.5);
  while (r * r > v) r--;
  while ((r + 1) * (r + 1) <= v) r++;
  return r;
}

void pix2ang_ring(long nside, long ipix, double* theta, double* phi) {
  const long nl2 = 2 * nside, nl4 = 4 * nside, npix = 12 * nside * nside;
  const long ncap = nl2 * (nside - 1);
  const double fact2 = 4.0 / npix;
  double z;
  if (ipix < ncap) {
    const long iring = (1 + isqrt_l(1 + 2 * ipix)) >> 1;
    const long iphi = ipix + 1 - 2 * iring * (iring - 1);
    z = 1.0 - (double)(iring * iring) * fact2;
    *phi = ((double)iphi - 0.5) * (0.5 * kPiArt) / iring;
  } else if (ipix < npix - ncap) {
    const long ip = ipix - ncap;
    const long iring = ip / nl4 + nside;
    const long iphi = ip % nl4 + 1;
    const double fodd = ((iring + nside) & 1) ? 1.0 : 0.5;
    z = (double)(nl2 - iring) * (2.0 / (3.0 * nside));
    *phi = ((double)iphi - fodd) * kPiArt / nl2;
  } else {
    const long ip = npix - ipix;
    const long iring = (1 + isqrt_l(2 * ip - 1)) >> 1;
    const long iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
    z = -1.0 + (double)(iring * iring) * fact2;
    *phi = ((double)iphi - 0.5) * (0.5 * kPiArt) / iring;
  }
  *theta = acos(z);
}

// set_angle_info + sort_angle_id (set_angle.c:12-247)
void make_angles(int nside, angle_info* out, int corner_num[8]) {
  const long npix = 12L * nside * nside;
  const double cos_pi_4 = cos(0.7853981633974483096156608458198757);
  for (int i = 0; i < 8; i++) corner_num[i] = 0;
  for (long ip = 0; ip < npix; ip++) {
    double th, ph;
    pix2ang_ring(nside, ip, &th, &ph);
    const double ct = cos(th), st = sin(th), cp = cos(ph), sp = sin(ph), tt = tan(th);
    double x = st * cp, y = st * sp, z = ct;
    if (fabs(x) < kTinyArt) x = x >= 0.0 ? kTinyArt : -kTinyArt;
    if (fabs(y) < kTinyArt) y = y >= 0.0 ? kTinyArt : -kTinyArt;
    if (fabs(z) < kTinyArt) z = z >= 0.0 ? kTinyArt : -kTinyArt;
    angle_info& a = out[ip];
    a.xovr = (float)x;
    a.yovr = (float)y;
    a.zovr = (float)z;
    const double rc = 1.0 / fabs(cp), rs = 1.0 / fabs(sp);
    int b;
    if (fabs(cp) >= cos_pi_4) {
      if (0.0 <= tt && tt <= rc) b = 0;
      else if (-rc <= tt && tt < 0.0) b = 1;
      else b = cp >= 0.0 ? 2 : 3;
    } else {
      if (0.0 <= tt && tt <= rs) b = 0;
      else if (-rs <= tt && tt < 0.0) b = 1;
      else b = sp >= 0.0 ? 4 : 5;
    }
    int cid = (cp > 0.0) ? (sp > 0.0 ? (ct > 0.0 ? 0 : 1) : (ct > 0.0 ? 2 : 3))
                         : (sp > 0.0 ? (ct > 0.0 ? 4 : 5) : (ct > 0.0 ? 6 : 7));
    a.base_id = (short)b;
    a.corner_id = (short)cid;
    corner_num[cid]++;
  }
  // sort_angle_id: bubble passes by corner, then by base within a corner
  for (long i = 0; i < npix - 1; i++)
    for (long j = npix - 1; j > i; j--)
      if (out[j - 1].corner_id > out[j].corner_id) std::swap(out[j - 1], out[j]);
  for (long i = 0; i < npix - 1; i++)
    for (long j = npix - 1; j > i; j--)
      if (out[j - 1].base_id > out[j].base_id && !(out[j - 1].corner_id < out[j].corner_id))
        std::swap(out[j - 1], out[j]);
}

double csect_HI(double nu) {
  if (nu < 1.0) return 0.0;
  const double eps = sqrt(nu - 1.0);
  return 6.3e-18 / (nu * nu * nu * nu) * exp(4.0 - 4.0 * atan(eps) / eps) /
         (1.0 - exp(-2.0 * kPiArt / eps));
}
double csect_HeI(double nu) {
  if (nu < 1.8125) return 0.0;
  return 7.83e-18 * (1.66 * pow((nu / 1.8125), -2.05) - 0.66 * pow((nu / 1.8125), -3.05));
}
double csect_HeII(double nu) {
  if (nu < 4.0) return 0.0;
  const double eps = sqrt(nu / 4.0 - 1.0);
  return 6.3e-18 * 256.0 / (nu * nu * nu * nu) * exp(4.0 - 4.0 * atan(eps) / eps) /
         (1.0 - exp(-2.0 * kPiArt / eps)) / 4.0;
}

// set_step_func_factor (set_step_func_factor.c:29-104)
void step_factors(int heatcool, step_func_factor* sf) {
  const double range = heatcool ? 0.0632353 : 0.0;
  const int nbin = heatcool ? 64 : 1;
  const double unit[3] = {1.0, 1.8125, 4.0};
  const double nuL[3] = {3.2885e15, 5.9483e15, 1.3154e16};
  double (*cs[3])(double) = {csect_HI, csect_HeI, csect_HeII};
  float* out[3] = {sf->HI, sf->HeI, sf->HeII};
  for (int s = 0; s < 3; s++) out[s][0] = out[s][1] = 0.0f;
  const double block = range / nbin;
  for (int inu = 0; inu < nbin; inu++)
    for (int s = 0; s < 3; s++) {
      double nu = unit[s] + inu * block, nu_next = unit[s] + (inu + 1) * block;
      if (nu == unit[s]) nu = unit[s] + 1e-7;
      if (nu_next == unit[s]) nu_next = unit[s] + 1e-7;
      const double c0 = 4.0e0 * kPiArt * cs[s](nu) / (kHplanck * (nu * nuL[s]));
      const double c0n = 4.0e0 * kPiArt * cs[s](nu_next) / (kHplanck * (nu_next * nuL[s]));
      const double c1 = inu == 0 ? 0.0e0 : c0 * (nu - unit[s]);
      const double c1n = c0 * (nu_next - unit[s]);
      out[s][0] += (c0 + c0n) * 0.5e0;
      out[s][1] += (c1 + c1n) * 0.5e0;
    }
  for (int s = 0; s < 3; s++) {
    out[s][0] /= nbin;
    out[s][1] /= nbin / (kHplanck * nuL[s]);
  }
}

}  // namespace
}  // namespace argot

using namespace argot;

extern "C" {

void argot_hip_set_diffuse_nside(int nside) { g_nside = nside > 0 ? nside : 8; }
void argot_hip_set_diffuse_heatcool(int on) { g_heatcool = on ? 1 : 0; }
void argot_hip_set_diffuse_batch(int ndir) { g_batch = ndir > 0 ? ndir : 0; }
void argot_hip_set_diffuse_kahan(int on) { g_kahan = on ? 1 : 0; }
void argot_hip_set_diffuse_exp64(int on) { g_exp64 = on ? 1 : 0; }
void argot_hip_diffuse_stats(struct argot_diffuse_stats* out) {
  if (out) *out = g_dstats;
}

void init_hd_param(struct host_diffuse_param* hd) {
  argot_hip_internal_clear_error();
  if (!hd) return argot_hip_internal_set_error(ARGOT_ERR_ARG);
  hd->step_fact = (struct step_func_factor*)malloc(sizeof(struct step_func_factor));
  hd->angle = (struct angle_info*)malloc(sizeof(struct angle_info) * 12 * g_nside * g_nside);
  if (!hd->step_fact || !hd->angle) argot_hip_internal_set_error(ARGOT_ERR_NOMEM);
}

void free_hd_param(struct host_diffuse_param* hd) {
  if (!hd) return;
  free(hd->step_fact);
  free(hd->angle);
  hd->step_fact = nullptr;
  hd->angle = nullptr;
}

void setup_diffuse_photon(struct host_diffuse_param* hd, struct run_param* this_run) {
  argot_hip_internal_clear_error();
  if (!hd || !hd->angle || !hd->step_fact || !this_run)
    return argot_hip_internal_set_error(ARGOT_ERR_ARG);
  make_angles(g_nside, hd->angle, hd->corner_id_num);
  step_factors(g_heatcool, hd->step_fact);
}

void init_diffuse_gpu(struct cuda_diffuse_param* cd, struct cuda_param* this_cuda,
                      struct run_param* this_run) {
  argot_hip_internal_clear_error();
  (void)this_cuda;
  if (!cd || !this_run) return argot_hip_internal_set_error(ARGOT_ERR_ARG);
  const int nx = this_run->nmesh_x_local, ny = this_run->nmesh_y_local,
            nz = this_run->nmesh_z_local;
  if (nx <= 0 || ny <= 0 || nz <= 0 || (nx | ny | nz) & 1)
    return argot_hip_internal_set_error(ARGOT_ERR_ARG);   // ray groups need even sizes
  const int64_t ncell = (int64_t)nx * ny * nz;
  const int nang = 12 * g_nside * g_nside;
  // Each direction of a batch owns one float4 slice over the mesh; more
  // directions per batch keep more rays in flight.
  int batch = g_batch;
  if (batch <= 0) {
    const int64_t fit = ((int64_t)8 << 30) / (16 * ncell);
    batch = (int)std::max<int64_t>(1, std::min<int64_t>(128, fit));
  }
  batch = std::min(batch, nang);
  if (!a_.stream) {
    A_OK(hipStreamCreate(&a_.stream));
    for (auto& e : a_.ev) A_OK(hipEventCreate(&e));
  }
  for (void* p : {(void*)a_.d_dirs, (void*)a_.med, (void*)a_.acc, (void*)a_.gh,
                  (void*)a_.comp, (void*)a_.ncross, (void*)a_.medT})
    if (p) A_OK(hipFree(p));
  a_.d_dirs = nullptr;
  a_.med = a_.acc = nullptr;
  a_.medT = nullptr;
  a_.gh = a_.comp = nullptr;
  a_.ncross = nullptr;
  a_.dirs.clear();
  // The sizes are set only once every buffer exists, so that a failed
  // allocation leaves a context that every later call refuses.
  a_.ncell = 0;
  a_.batch = 0;
  A_OK(hipMalloc(&a_.d_dirs, sizeof(ArtDir) * nang));
  A_OK(hipMalloc(&a_.med, sizeof(float4) * 2 * ncell));
  A_OK(hipMalloc(&a_.acc, sizeof(float4) * ncell * batch));
  A_OK(hipMalloc(&a_.medT, sizeof(float4) * 2 * ncell));
  A_OK(hipMalloc(&a_.gh, sizeof(float) * 6 * ncell));
  A_OK(hipMalloc(&a_.comp, sizeof(float) * 6 * ncell));
  A_OK(hipMalloc(&a_.ncross, sizeof(unsigned long long)));
  a_.ncell = ncell;
  a_.batch = batch;
  a_.nside = g_nside;
  // the reference's pointers: step factors and angles live in the
  // library's context; rmesh names the accumulator slices
  cd[0].step_fact = nullptr;
  cd[0].angle = nullptr;
  cd[0].rmesh = reinterpret_cast<struct radiation_mesh*>(a_.acc);
}

void send_diffuse_data(struct host_diffuse_param* hd, struct cuda_diffuse_param* cd,
                       struct cuda_param* this_cuda) {
  argot_hip_internal_clear_error();
  (void)cd;
  (void)this_cuda;
  if (!hd || !hd->angle || !hd->step_fact) return argot_hip_internal_set_error(ARGOT_ERR_ARG);
  if (!a_.d_dirs || a_.ncell <= 0) return argot_hip_internal_set_error(ARGOT_ERR_STATE);
  const int nang = 12 * a_.nside * a_.nside;
  a_.dirs.resize(nang);
  for (int i = 0; i < nang; i++) {
    const angle_info& an = hd->angle[i];
    ArtDir& d = a_.dirs[i];
    d.ovr[0] = an.xovr;
    d.ovr[1] = an.yovr;
    d.ovr[2] = an.zovr;
    d.base = an.base_id;
    for (int k = 0; k < 3; k++) d.rcp[k] = 1.0 / (double)d.ovr[k];
  }
  const step_func_factor* sf = hd->step_fact;
  const float f[6] = {sf->HI[0], sf->HI[1], sf->HeI[0], sf->HeI[1], sf->HeII[0], sf->HeII[1]};
  memcpy(a_.fact, f, sizeof f);
  A_OK(hipMemcpy(a_.d_dirs, a_.dirs.data(), sizeof(ArtDir) * nang, hipMemcpyHostToDevice));
}

void calc_diffuse_photon_radiation(struct fluid_mesh* mesh, struct run_param* this_run,
                                   struct cuda_mem_space* cuda_mem, struct cuda_param* this_cuda,
                                   struct host_diffuse_param* hd, struct cuda_diffuse_param* cd,
                                   struct omp_param* this_omp) {
  argot_hip_internal_clear_error();
  (void)mesh;
  (void)this_cuda;
  (void)hd;
  (void)cd;
  (void)this_omp;
  const double t_start = now_ms();
  if (!this_run || !cuda_mem || !cuda_mem->mesh_dev)
    return argot_hip_internal_set_error(ARGOT_ERR_STATE);
  if (a_.dirs.empty() || !a_.d_dirs || !a_.med || !a_.acc || !a_.medT || !a_.gh || !a_.comp ||
      !a_.ncross || a_.ncell <= 0 || a_.batch <= 0)
    return argot_hip_internal_set_error(ARGOT_ERR_STATE);
  ArtGrid g;
  g.nx = this_run->nmesh_x_local;
  g.ny = this_run->nmesh_y_local;
  g.nz = this_run->nmesh_z_local;
  g.ncell = (int64_t)g.nx * g.ny * g.nz;
  if (g.ncell != a_.ncell || this_run->rank_x || this_run->rank_y || this_run->rank_z ||
      this_run->nnode_x > 1 || this_run->nnode_y > 1 || this_run->nnode_z > 1)
    return argot_hip_internal_set_error(ARGOT_ERR_ARG);
  if (this_run->delta_x != this_run->delta_y || this_run->delta_x != this_run->delta_z)
    return argot_hip_internal_set_error(ARGOT_ERR_ARG);   // the ray groups assume cubic cells
  g.xmin = this_run->xmin_local;
  g.ymin = this_run->ymin_local;
  g.zmin = this_run->zmin_local;
  g.xmax = this_run->xmax_local;
  g.ymax = this_run->ymax_local;
  g.zmax = this_run->zmax_local;
  g.dx = this_run->delta_x;
  g.dy = this_run->delta_y;
  g.dz = this_run->delta_z;
  g.lunit = this_run->lunit;
  unsigned char* dmesh = reinterpret_cast<unsigned char*>(cuda_mem->mesh_dev);
  hipStream_t st = a_.stream;
  const int64_t nc = g.ncell;
  const unsigned nb = (unsigned)((nc + 255) / 256);
  A_OK(hipEventRecord(a_.ev[0], st));
  art_rmesh_kernel<<<nb, 256, 0, st>>>(
      dmesh, g.nx, g.ny, g.nz, this_run->denstonh, this_run->uenetok,
      argot_hip_internal_helium_fact(), (float)csect_HI(1.0 + 1.0e-7),
      (float)csect_HeI(1.8125 + 1.0e-7), (float)csect_HeII(4.0 + 1.0e-7), a_.med, a_.medT);
  A_OK(hipGetLastError());
  A_OK(hipMemsetAsync(a_.gh, 0, sizeof(float) * 6 * nc, st));
  A_OK(hipMemsetAsync(a_.comp, 0, sizeof(float) * 6 * nc, st));
  A_OK(hipMemsetAsync(a_.ncross, 0, sizeof(unsigned long long), st));
  const int nang = (int)a_.dirs.size();
  const int nray_max = std::max(g.nx * g.ny, std::max(g.ny * g.nz, g.nz * g.nx)) / 4;
  float trace_ms = 0.0f, gh_ms = 0.0f;
  for (int d0 = 0; d0 < nang; d0 += a_.batch) {
    const int nd = std::min(a_.batch, nang - d0);
    const int64_t nacc = (int64_t)nd * nc;
    art_fill_kernel<<<(unsigned)((nacc + 255) / 256), 256, 0, st>>>(a_.acc, nacc);
    A_OK(hipEventRecord(a_.ev[1], st));
    TraceArgs ta;
    ta.g = g;
    ta.dirs = a_.d_dirs + d0;
    ta.ndir = nd;
    ta.nray_max = nray_max;
    ta.med = a_.med;
    ta.medT = a_.medT;
    ta.acc = a_.acc;
    ta.ncross = a_.ncross;
    const int64_t nthread = (int64_t)nd * nray_max;
    for (int grp = 0; grp < 4; grp++) {
      ta.group = grp;
      if (g_exp64)
        art_trace_kernel<true><<<(unsigned)((nthread + 255) / 256), 256, 0, st>>>(ta);
      else
        art_trace_kernel<false><<<(unsigned)((nthread + 255) / 256), 256, 0, st>>>(ta);
    }
    A_OK(hipEventRecord(a_.ev[2], st));
    const float* f = a_.fact;
    const dim3 ghgrid((unsigned)(((g.ny + kGhTile - 1) / kGhTile) * ((g.nz + kGhTile - 1) / kGhTile)),
                      (unsigned)g.nx);
    if (g_kahan)
      art_gh_kernel<true><<<ghgrid, kGhTile * kGhTile, 0, st>>>(
          a_.acc, a_.d_dirs + d0, nd, g.nx, g.ny, g.nz, a_.med, f[0], f[1], f[2], f[3], f[4],
          f[5], a_.gh, a_.comp);
    else
      art_gh_kernel<false><<<ghgrid, kGhTile * kGhTile, 0, st>>>(
          a_.acc, a_.d_dirs + d0, nd, g.nx, g.ny, g.nz, a_.med, f[0], f[1], f[2], f[3], f[4],
          f[5], a_.gh, a_.comp);
    A_OK(hipGetLastError());
    A_OK(hipEventRecord(a_.ev[3], st));
    A_OK(hipEventSynchronize(a_.ev[3]));
    float m0 = 0.0f, m1 = 0.0f;
    A_OK(hipEventElapsedTime(&m0, a_.ev[1], a_.ev[2]));
    A_OK(hipEventElapsedTime(&m1, a_.ev[2], a_.ev[3]));
    trace_ms += m0;
    gh_ms += m1;
  }
  const float r_n_ang = (float)(1.0e0 / nang);
  art_gh_sum_kernel<<<nb, 256, 0, st>>>(dmesh, nc, a_.gh, r_n_ang);
  A_OK(hipGetLastError());
  unsigned long long ncross = 0;
  A_OK(hipMemcpyAsync(&ncross, a_.ncross, sizeof ncross, hipMemcpyDeviceToHost, st));
  A_OK(hipStreamSynchronize(st));
  g_dstats.ndir = (uint64_t)nang;
  g_dstats.ncross = ncross;
  g_dstats.ms_trace = trace_ms;
  g_dstats.ms_gh = gh_ms;
  g_dstats.ms_total = now_ms() - t_start;
}

void free_diffuse_gpu(struct cuda_diffuse_param* cd, struct cuda_param* this_cuda) {
  (void)this_cuda;
  for (void* p : {(void*)a_.d_dirs, (void*)a_.med, (void*)a_.acc, (void*)a_.gh,
                  (void*)a_.comp, (void*)a_.ncross, (void*)a_.medT})
    if (p) (void)hipFree(p);
  a_.d_dirs = nullptr;
  a_.med = a_.acc = nullptr;
  a_.medT = nullptr;
  a_.gh = a_.comp = nullptr;
  a_.ncross = nullptr;
  a_.dirs.clear();
  a_.ncell = 0;
  a_.batch = 0;
  if (cd) cd[0].rmesh = nullptr;
}

}  // extern "C"
