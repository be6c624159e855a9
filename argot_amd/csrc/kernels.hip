// HIP kernels of the point-source radiation step for gfx950 (MI355X).
//
//   pack_fields         AoS fluid_mesh -> per-cell FP32 (nHI, nHeI, nHeII)
//   zero_rates          prev_chem rates := 0         (zero_out_photoion_rate)
//   add_rates           prev_chem rates += table     (merge_photoion_data)
//   walk_segments       optical depth of a segment list (calc_optical_depth)
//   target_rates        per-target rates and sums of a ray list (calc_photoion_rate)
//   fused_step          tree walk + segment + optical depth + rates + per-cell
//                       sum for 4x4x4 target tiles, one wavefront per tile
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "argot_device.h"
#include "argot_internal.h"
#include "kernels.h"

namespace argot {

// Offsets inside struct fluid_mesh (156 B, include/argot_hip.h).
constexpr int kMeshStride = 156;
constexpr int kOffDens = 0;
constexpr int kOffFHI = 52;
constexpr int kOffFHeI = 60;
constexpr int kOffFHeII = 64;
constexpr int kOffRates = 124;   // prev_chem.GammaHI .. HeatHeII (6 floats)

__device__ __forceinline__ const float* mesh_f(const unsigned char* mesh, int64_t c, int off) {
  return reinterpret_cast<const float*>(mesh + c * kMeshStride + off);
}
__device__ __forceinline__ float* mesh_fw(unsigned char* mesh, int64_t c, int off) {
  return reinterpret_cast<float*>(mesh + c * kMeshStride + off);
}

// calc_optical_depth.c:150-156: nHI = dens*fHI, nHe = dens*HELIUM_FACT in
// double rounded to float, nHeI = nHe*fHeI, all FP32.
__device__ __forceinline__ float4 form_fields(const unsigned char* __restrict__ mesh, int64_t c,
                                              double helium_fact) {
  const float nH = *mesh_f(mesh, c, kOffDens);
  const float nHI = nH * *mesh_f(mesh, c, kOffFHI);
  const float nHe = (float)((double)nH * helium_fact);
  const float nHeI = nHe * *mesh_f(mesh, c, kOffFHeI);
  const float nHeII = nHe * *mesh_f(mesh, c, kOffFHeII);
  return make_float4(nHI, nHeI, nHeII, 0.0f);
}

__global__ void pack_fields_kernel(const unsigned char* __restrict__ mesh, int64_t ncell,
                                   double helium_fact, float4* __restrict__ nf) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  nf[c] = form_fields(mesh, c, helium_fact);
}

// Per-cell densities formed from the AoS mesh records at each crossing, with
// pack_fields_kernel's arithmetic (so the optical depths are the packed
// walk's bits): the walk of a call that has no packed-field buffer.
struct MeshFields {
  const unsigned char* __restrict__ mesh;
  double helium_fact;
  static constexpr uint32_t kScale = 1u;
  static constexpr bool kChecked = false;
  __device__ __forceinline__ float4 operator()(uint32_t idx) const {
    return form_fields(mesh, (int64_t)idx, helium_fact);
  }
};

__global__ void zero_rates_kernel(unsigned char* __restrict__ mesh, int64_t ncell) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  float* r = mesh_fw(mesh, c, kOffRates);
#pragma unroll
  for (int k = 0; k < 6; k++) r[k] = 0.0f;
}

__global__ void add_rates_kernel(unsigned char* __restrict__ mesh, const float* __restrict__ rate,
                                 int64_t ncell) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  float* r = mesh_fw(mesh, c, kOffRates);
#pragma unroll
  for (int k = 0; k < 6; k++) r[k] = r[k] + rate[6 * c + k];
}

// The six rates of every cell between the mesh and a photoion_rate table
// (24 B per cell, the prev_chem rates in the same order):
// recv_photoion_data_kernel / send_photoion_data_kernel of
// src/recv_photoion_data.cu and src/send_photoion_data.cu.
__global__ void get_rates_kernel(const unsigned char* __restrict__ mesh, int64_t ncell,
                                 float* __restrict__ rate) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const float* r = mesh_f(mesh, c, kOffRates);
#pragma unroll
  for (int k = 0; k < 6; k++) rate[6 * c + k] = r[k];
}

__global__ void set_rates_kernel(unsigned char* __restrict__ mesh, const float* __restrict__ rate,
                                 int64_t ncell) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  float* r = mesh_fw(mesh, c, kOffRates);
#pragma unroll
  for (int k = 0; k < 6; k++) r[k] = rate[6 * c + k];
}

// ray_segment records (56 B): start @16, end @28, tau @40.  A block takes
// 256 consecutive segments and hands them to its lanes in order of their
// estimated crossing count (counting sort in LDS), so that the lanes of a
// wavefront walk segments of similar length; each segment's result does not
// depend on which lane walks it.
// Fields: where the per-cell densities come from (PackedFields: the packed
// float4 array; MeshFields: the mesh records).  POW2: the DDA's face-position
// form for power-of-two cell sizes.
constexpr int kWalkBlock = 256;
template <bool POW2, typename Fields>
__global__ __launch_bounds__(kWalkBlock) void walk_segments_kernel(
    GridDev g, Fields fields, unsigned char* __restrict__ seg, int64_t nseg, int kahan) {
  __shared__ unsigned bucket_lds[64];
  __shared__ unsigned perm_lds[kWalkBlock];
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kWalkBlock;
  if (t < 64) bucket_lds[t] = 0;
  __syncthreads();
  unsigned b = 0;
  if (base + t < nseg) {
    const float* r = reinterpret_cast<const float*>(seg + (base + t) * 56);
    const float est = fabsf(r[7] - r[4]) / g.dx + fabsf(r[8] - r[5]) / g.dy +
                      fabsf(r[9] - r[6]) / g.dz;
    b = est < 504.0f ? (unsigned)est >> 3 : 63u;   // buckets of 8 crossings
  }
  const unsigned pos = atomicAdd(&bucket_lds[b], 1u);
  __syncthreads();
  if (t < 64) {   // exclusive prefix sum of the 64 bucket sizes (one wavefront)
    const unsigned v = bucket_lds[t];
    unsigned incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned u = __shfl_up(incl, d, 64);
      if (t >= d) incl += u;
    }
    bucket_lds[t] = incl - v;
  }
  __syncthreads();
  perm_lds[bucket_lds[b] + pos] = (unsigned)t;
  __syncthreads();
  const int64_t i = base + perm_lds[t];
  if (i >= nseg) return;
  float* rec = reinterpret_cast<float*>(seg + i * 56);
  const Vec3f s{rec[4], rec[5], rec[6]};
  const Vec3f e{rec[7], rec[8], rec[9]};
  WalkOut w;
  if (kahan)
    walk_segment_with<true, false, POW2>(g, fields, s, e, &w);
  else
    walk_segment_with<false, false, POW2>(g, fields, s, e, &w);
  rec[10] = w.tau[0];
  rec[11] = w.tau[1];
  rec[12] = w.tau[2];
}

// Bins [k0, k1) of the FP32 frequency loop with NS active species (the
// species s >= NS have sigma_s = 0 in these bins, so their optical depth
// enters only through the constant cst).  Tables are read with scalar loads
// (wave-uniform bin index); sums run in FP32 over 32 bins, then in FP64.
constexpr float kLog2e = 1.4426950408889634f;

// sinh(y)/y for 0 <= y <= 2.5 from yl = -log2(e) y: a degree-4 polynomial
// in u = y^2 (minimax relative error over u in [0, 6.25]: 3.3e-7 before
// rounding, 4.4e-7 after FP32 evaluation), evaluated in yl^2 with
// coefficient k scaled by log2(e)^(-2k).
constexpr double kL2 = 1.4426950408889634 * 1.4426950408889634;
constexpr float kSh0 = (float)1.0000003313e+00;
constexpr float kSh1 = (float)(1.6666356469e-01 / kL2);
constexpr float kSh2 = (float)(8.3377321051e-03 / (kL2 * kL2));
constexpr float kSh3 = (float)(1.9630612284e-04 / (kL2 * kL2 * kL2));
constexpr float kSh4 = (float)(3.1551783390e-06 / (kL2 * kL2 * kL2 * kL2));
// The polynomial's range: in-cell optical depth 2y <= kThinMax, i.e.
// yl >= kYlThin.
constexpr double kThinMax = 5.0;
constexpr float kYlThin = (float)(-0.5 * kThinMax * 1.4426950408889634);
__device__ __forceinline__ float sinhc_l(float yl) {
  const float u = yl * yl;
  float p = __builtin_fmaf(u, kSh4, kSh3);
  p = __builtin_fmaf(u, p, kSh2);
  p = __builtin_fmaf(u, p, kSh1);
  return __builtin_fmaf(u, p, kSh0);
}

// Bins [k0, k1) of a ray whose in-cell optical depth is at most kThinMax in
// every bin: exp(-tau) (1 - exp(-dtau)) / dtau = exp(-(tau + y)) sinh(y)/y with
// y = dtau/2, one exponential per bin.  dl holds the ray's in-cell weights
// times -log2(e)/2, so that their sum with the cross sections is
// yl = -log2(e) y and 2^(yl - log2(e) tau) = exp(-(tau + y)).  FOLD (bins
// where cst = 0): nl holds the column densities times -log2(e), and the
// exponent is summed onto yl directly, one instruction fewer per bin.
template <int NS, bool FOLD, typename SpecF>
__device__ __forceinline__ void rate_bins_thin(const RateTables& rt, int k0, int k1, float cst,
                                               const float n[3], const float nl[3],
                                               const float dl[3], SpecF spec, double G[3],
                                               double H[3]) {
  const int ng = rt.ngrid;
  for (int c = k0; c < k1; c += 32) {
    float g[NS], h[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) g[s] = h[s] = 0.0f;
    const int ce = min(c + 32, k1);
#pragma unroll 4
    for (int k = c; k < ce; k++) {
      float sg[NS], hs[NS];
      float yl = 0.0f, z;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        sg[s] = ldc(rt.sig + s * ng + k);
        hs[s] = ldc(rt.hsig + s * ng + k);
        yl = __builtin_fmaf(dl[s], sg[s], yl);
      }
      if constexpr (FOLD) {
        z = yl;
#pragma unroll
        for (int s = 0; s < NS; s++) z = __builtin_fmaf(nl[s], ldc(rt.dsig + s * ng + k), z);
      } else {
        float tau = cst;
#pragma unroll
        for (int s = 0; s < NS; s++) tau = __builtin_fmaf(n[s], ldc(rt.dsig + s * ng + k), tau);
        z = __builtin_fmaf(tau, -kLog2e, yl);
      }
      const float e = __builtin_amdgcn_exp2f(z);
      const float b = spec(k) * e * sinhc_l(yl);
#pragma unroll
      for (int s = 0; s < NS; s++) {
        g[s] = __builtin_fmaf(sg[s], b, g[s]);
        h[s] = __builtin_fmaf(hs[s], b, h[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      G[s] += g[s];
      H[s] += h[s];
    }
  }
}

// Bins [k0, k1) for a wavefront in which some lane's ray has an in-cell
// optical depth above kThinMax.  A lane flagged thin computes exactly what
// rate_bins_thin computes (the same operations in the same order), so a
// cell's rates do not depend on which rays share its wavefront.  The other
// lanes take, bin by bin, the same thin form where yl >= kYlThin, and
// otherwise exp(-tau) (1 - exp(-2y)) / (2y) with the optical depth tau
// summed on its own (no cancellation: exp(-2y) <= exp(-kThinMax)).
template <int NS, bool FOLD, typename SpecF>
__device__ __forceinline__ void rate_bins_mixed(const RateTables& rt, int k0, int k1, float cst,
                                                const float n[3], const float nl[3],
                                                const float dl[3], bool thin, SpecF spec,
                                                double G[3], double H[3]) {
  constexpr float kHalfLog2eN = -0.5f * kLog2e;  // 1/(2y) = kHalfLog2eN / yl
  const int ng = rt.ngrid;
  for (int c = k0; c < k1; c += 32) {
    float g[NS], h[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) g[s] = h[s] = 0.0f;
    const int ce = min(c + 32, k1);
#pragma unroll 2
    for (int k = c; k < ce; k++) {
      float sg[NS], hs[NS], ds[NS];
      float yl = 0.0f, z, tau = cst;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        sg[s] = ldc(rt.sig + s * ng + k);
        hs[s] = ldc(rt.hsig + s * ng + k);
        ds[s] = ldc(rt.dsig + s * ng + k);
        yl = __builtin_fmaf(dl[s], sg[s], yl);
      }
#pragma unroll
      for (int s = 0; s < NS; s++) tau = __builtin_fmaf(n[s], ds[s], tau);
      if constexpr (FOLD) {
        z = yl;
#pragma unroll
        for (int s = 0; s < NS; s++) z = __builtin_fmaf(nl[s], ds[s], z);
      } else {
        z = __builtin_fmaf(tau, -kLog2e, yl);
      }
      const float sk = spec(k);
      const float bt = sk * __builtin_amdgcn_exp2f(z) * sinhc_l(yl);
      const float e1 = __builtin_amdgcn_exp2f(tau * -kLog2e);
      const float om = 1.0f - __builtin_amdgcn_exp2f(yl + yl);
      const float bk = sk * e1 * om * (__builtin_amdgcn_rcpf(yl) * kHalfLog2eN);
      const float b = (thin || yl >= kYlThin) ? bt : bk;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        g[s] = __builtin_fmaf(sg[s], b, g[s]);
        h[s] = __builtin_fmaf(hs[s], b, h[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      G[s] += g[s];
      H[s] += h[s];
    }
  }
}

// FP32 frequency loop (photon-conserving form, calc_photoion_rate.c:327-386
// regrouped): with B_k = N_k exp(-tau_k) (1 - exp(-dtau_k)) / dtau_k,
//   Gamma_s = r_s d_s sum_k sigma_s(k) B_k,  Heat_s = r_s d_s sum_k hsig_s(k) B_k.
// The optical depth is shifted by its last-bin value tsh (folded back in
// FP64): tau_k - tsh = sum_s N_s (sigma_s(k) - sigma_s(last)).  Bins below
// the HeI and HeII edges skip the absent species.
template <typename SpecF>
__device__ inline void rate_loop_f32(const RateTables& rt, const RayRate& rr, double scale,
                                     SpecF spec, double out[6]) {
  const double tsh = rr.t[0] * rt.cs_last[0] + rr.t[1] * rt.cs_last[1] + rr.t[2] * rt.cs_last[2];
  const float n[3] = {(float)rr.t[0], (float)rr.t[1], (float)rr.t[2]};
  const float c1 = (float)(-(rr.t[1] * rt.cs_last[1] + rr.t[2] * rt.cs_last[2]));
  const float c2 = (float)(-(rr.t[2] * rt.cs_last[2]));
  // per-ray prefactors first (same operations as after the loop), so that
  // only they, not the ray's FP64 quantities, stay live through the bins
  // exp(-tsh) as 2^n 2^f, n = floor(-tsh log2 e) and f in [0, 1): the FP32
  // exponential of f (relative error about 1e-7, below the loop's own) times
  // an exact power of two, in FP64 range.  The library exponential's
  // polynomial would hold a dozen FP64 constants in registers across the
  // kernel, which the register budget spills.
  constexpr double kLog2eP = 1.4426950408889634;
  const double xl = -tsh * kLog2eP;
  const double fl = fmax(floor(xl), -2000.0);
  const double pre =
      scale * ldexp((double)__builtin_amdgcn_exp2f((float)(xl - fl)), (int)fl);
  const double p0 = rr.r[0] * rr.d[0] * pre;
  const double p1 = rr.r[1] * rr.d[1] * pre;
  const double p2 = rr.r[2] * rr.d[2] * pre;
  double G[3] = {0, 0, 0}, H[3] = {0, 0, 0};
  // largest in-cell optical depth of this ray over the bins.  A wavefront
  // whose rays are all thin runs the thin loop; one with a thicker ray runs
  // the combined loop, in which a thin lane does the thin loop's operations
  // exactly, so a cell's rates do not depend on which rays share its
  // wavefront (tile shape, grid size).
  const double dmax = rr.d[0] * rt.cs_max[0] + rr.d[1] * rt.cs_max[1] + rr.d[2] * rt.cs_max[2];
  constexpr double kHalfLog2e = 0.5 * 1.4426950408889634;
  constexpr double kLog2eD = 1.4426950408889634;
  const float dl[3] = {(float)(rr.d[0] * -kHalfLog2e), (float)(rr.d[1] * -kHalfLog2e),
                       (float)(rr.d[2] * -kHalfLog2e)};
  const float nl[3] = {(float)(rr.t[0] * -kLog2eD), (float)(rr.t[1] * -kLog2eD),
                       (float)(rr.t[2] * -kLog2eD)};
  const bool thin = dmax <= kThinMax;
  if (__all(thin)) {
    rate_bins_thin<1, false>(rt, 0, rt.k_he1, c1, n, nl, dl, spec, G, H);
    rate_bins_thin<2, false>(rt, rt.k_he1, rt.k_he2, c2, n, nl, dl, spec, G, H);
    rate_bins_thin<3, true>(rt, rt.k_he2, rt.ngrid, 0.0f, n, nl, dl, spec, G, H);
  } else {
    rate_bins_mixed<1, false>(rt, 0, rt.k_he1, c1, n, nl, dl, thin, spec, G, H);
    rate_bins_mixed<2, false>(rt, rt.k_he1, rt.k_he2, c2, n, nl, dl, thin, spec, G, H);
    rate_bins_mixed<3, true>(rt, rt.k_he2, rt.ngrid, 0.0f, n, nl, dl, thin, spec, G, H);
  }
  out[0] = p0 * G[0];
  out[1] = p0 * H[0] * kHnuL;
  out[2] = p1 * G[1];
  out[3] = p2 * G[2];
  out[4] = p1 * H[1] * kHnuL;
  out[5] = p2 * H[2] * kHnuL;
}

// Lower bound of the shifted bin optical depth and upper bound of the
// in-cell optical depth decide whether the FP32 loop is safe.
__device__ __forceinline__ bool fast_rate_ok(const RateTables& rt, const RayRate& rr) {
  double lo = 0.0, dmax = 0.0;
  bool finite = true;
  for (int s = 0; s < 3; s++) {
    lo += rr.t[s] >= 0.0 ? -rr.t[s] * rt.cs_last[s] : rr.t[s] * (rt.cs_max[s] - rt.cs_last[s]);
    dmax += rr.d[s] * rt.cs_max[s];
    finite = finite && fabs(rr.t[s]) < 1.0e36 && rr.d[s] < 1.0e36;
  }
  return finite && lo > -60.0 && dmax < 1.0e30;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ARGOT_WAVES_PER_EU (build option, 6 by default) caps the register budget
// of the exact-rate fused kernels so that this many wavefronts fit per SIMD:
// their walks wait on gathers, and more wavefronts hide more of that wait
// than the spilled registers cost (on 256^3 / 4096 sources, 140 ms at 6
// against 143 ms at 5, 148 ms at 4 and 149 ms at 8).  The surrogate variant,
// whose LDS holds about 4 wavefronts per SIMD, and the per-call rate kernel,
// which does not walk, keep the budget of 4.
#if defined(ARGOT_WAVES_PER_EU) && ARGOT_WAVES_PER_EU > 0
#define ARGOT_OCCUPANCY_FOR(SURR) \
  __attribute__((amdgpu_waves_per_eu((SURR) ? 4 : ARGOT_WAVES_PER_EU)))
#else
#define ARGOT_OCCUPANCY_FOR(SURR)
#endif
#define ARGOT_CALL_OCCUPANCY __attribute__((amdgpu_waves_per_eu(4)))

// calc_photoion_rate_cuda (calc_photoion_rate_cuda.cu:92-216 with the
// per-target merge of calc_photoion_rate.c:403-425): one lane per target
// cell.  A target's rays are consecutive in the caller's order and its lane
// sums their rates in FP64 in that order, then adds the sum to the cell
// once: no atomics, no per-ray buffer.  Each ray names one of the call's
// distinct spectra (rays from one source or tree node share it).  The
// wavefront visits the spectra its lanes' next rays use in increasing index
// order (the host numbers them by first appearance, so the lists of
// neighbouring targets run nearly in step): the lanes at that spectrum take
// their ray with the bins read through scalar loads, the others wait.  Each
// ray's arithmetic is the fused kernel's, so a ray list taken from the
// fused step gives its bits.
__global__ __launch_bounds__(64) ARGOT_CALL_OCCUPANCY void target_rates_kernel(
    GridDev g, RateTables rt, unsigned char* __restrict__ mesh, const CallTarget* __restrict__ tg,
    int64_t ntarget, const CallRay* __restrict__ rays, const float* __restrict__ spec32,
    const double* __restrict__ spec64, const double* __restrict__ spec_scale) {
  const int lane = threadIdx.x;
  const int64_t t = (int64_t)blockIdx.x * 64 + lane;
  const bool valid = t < ntarget;
  CallTarget T{0, 0, 0, {0, 0, 0}};
  if (valid) T = tg[t];
  float dens = 0.0f, fHI = 0.0f, fHeI = 0.0f, fHeII = 0.0f;
  if (valid) {
    dens = *mesh_f(mesh, T.cell, kOffDens);
    fHI = *mesh_f(mesh, T.cell, kOffFHI);
    fHeI = *mesh_f(mesh, T.cell, kOffFHeI);
    fHeII = *mesh_f(mesh, T.cell, kOffFHeII);
  }
  const Vec3f tgt{cell_centre(g.gxmin, T.g[0], g.dx), cell_centre(g.gymin, T.g[1], g.dy),
                  cell_centre(g.gzmin, T.g[2], g.dz)};
  int64_t r = T.first;
  const int64_t rend = T.first + T.count;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int ng = rt.ngrid;
  for (;;) {
    const bool more = r < rend;
    const int sp = more ? rays[r].spec : INT_MAX;
    const int u = wave_min(sp);
    if (u == INT_MAX) break;
    if (more && sp == u) {
      const CallRay R = rays[r];
      const Vec3f src{R.src[0], R.src[1], R.src[2]};
      RayRate rr;
      ray_rate_setup(g, rt, R.tau, src, tgt, dens, fHI, fHeI, fHeII, &rr);
      double res[6];
      if (fast_rate_ok(rt, rr)) {
        const float* s32 = spec32 + (int64_t)u * ng;
        rate_loop_f32(rt, rr, ldc(spec_scale + u), [&](int k) { return ldc(s32 + k); }, res);
      } else {
        const double* s64 = spec64 + (int64_t)u * ng;
        rate_loop_f64(rt, rr, [&](int k) { return ldc(s64 + k); }, res);
      }
#pragma unroll
      for (int k = 0; k < 6; k++) acc[k] += res[k];
      r++;
    }
  }
  if (valid && T.count > 0) {
    float* o = mesh_fw(mesh, T.cell, kOffRates);
#pragma unroll
    for (int k = 0; k < 6; k++) o[k] = (float)((double)o[k] + acc[k]);
  }
}

// Every third bit of v (bits 0, 3, 6, ...) packed together: one axis of a
// Morton index.
__device__ __forceinline__ uint32_t compact3(uint32_t v) {
  v &= 0x09249249u;
  v = (v ^ (v >> 2)) & 0x030c30c3u;
  v = (v ^ (v >> 4)) & 0x0300f00fu;
  v = (v ^ (v >> 8)) & 0xff0000ffu;
  v = (v ^ (v >> 16)) & 0x000003ffu;
  return v;
}

// Next tile for this wavefront: its own XCD's queue first (blocks are
// dispatched round-robin over the 8 XCDs), then the others in turn.
__device__ __forceinline__ int64_t next_tile(StepCounters* ctr, int64_t ntile, int* q, int* tried) {
  while (*tried < kNumQueues) {
    const int64_t lo = ntile * (*q) / kNumQueues;
    const int64_t hi = ntile * (*q + 1) / kNumQueues;
    int t = 0;
    if (threadIdx.x == 0) t = atomicAdd(&ctr->queue[*q], 1);
    t = __shfl(t, 0, 64);
    if (lo + t < hi) return lo + t;
    *q = (*q + 1) % kNumQueues;
    (*tried)++;
  }
  return -1;
}


// 2-point Gauss-Legendre nodes on [0, 1].
constexpr double kGl1 = 0.21132486540518711775, kGl2 = 0.78867513459481288225;

// Can the network serve this ray?  Its node's spectrum must have the
// trained shape (q > 0), the in-cell optical depth must be small enough
// for the 2-point rule, and both nodes of the in-cell path (threshold
// optical depths (t_s + x d_s) sig_th_s) must lie below the hull's upper
// edge and in trusted cells.  Below the lower edge the kernels are flat
// (relative change < 1e-5), so inputs are clamped there.
__device__ inline bool surr_inputs(const SurrDev& S, const RateTables& rt, const RayRate& rr,
                                   double q, float xa[3], float xb[3]) {
  if (!(q > 0.0)) return false;
  bool ok = true;
  double dmax = 0.0;
  int ca[3], cb[3];
#pragma unroll
  for (int s = 0; s < 3; s++) {
    const double t = rr.t[s], d = rr.d[s];
    ok = ok && t >= 0.0 && d >= 0.0;
    dmax += d * rt.cs_max[s];
    const double ta = (t + kGl1 * d) * (double)S.sig_th[s];
    const double tb = (t + kGl2 * d) * (double)S.sig_th[s];
    ok = ok && tb <= S.tau_hi;
    xa[s] = __log10f((float)fmax(ta, S.tau_lo));
    xb[s] = __log10f((float)fmax(tb, S.tau_lo));
    ca[s] = min(S.trust_n - 1, max(0, (int)((xa[s] - S.lo) * S.inv_cell)));
    cb[s] = min(S.trust_n - 1, max(0, (int)((xb[s] - S.lo) * S.inv_cell)));
  }
  ok = ok && dmax <= (double)S.dtau_max;
  if (!ok) return false;
  const int n = S.trust_n;
  return S.trust[(ca[0] * n + ca[1]) * n + ca[2]] && S.trust[(cb[0] * n + cb[1]) * n + cb[2]];
}

// 10^y in FP64 from an FP32 exponent (kernels reach 1e-90, below FP32).
__device__ __forceinline__ double exp10_wide(float y) {
  const double t = (double)y * 3.3219280948873623479;   // log2(10)
  const double n = floor(t);
  return ldexp((double)exp2f((float)(t - n)), (int)n);
}

// Rates of a ray from the network: Gamma_s = r_s d_s q <F_s>, Heat_s =
// r_s d_s q <G_s> hplanck nu_L, with <.> the 2-point mean over the in-cell
// path (the photon-conserving cell average of exp(-tau)).
__device__ inline void surr_rates(const RayRate& rr, double q, const float ya[kNout],
                                  const float yb[kNout], double res[6]) {
  double m[kNout];
#pragma unroll
  for (int o = 0; o < kNout; o++) m[o] = 0.5 * (exp10_wide(ya[o]) + exp10_wide(yb[o]));
  const double p0 = rr.r[0] * rr.d[0] * q, p1 = rr.r[1] * rr.d[1] * q, p2 = rr.r[2] * rr.d[2] * q;
  res[0] = p0 * m[0];
  res[1] = p0 * m[3] * kHnuL;
  res[2] = p1 * m[1];
  res[3] = p2 * m[2];
  res[4] = p1 * m[4] * kHnuL;
  res[5] = p2 * m[5] * kHnuL;
}

template <bool KAHAN, bool CAPTURE, bool SURR, bool NARROW, bool POW2>
__global__ __launch_bounds__(64) ARGOT_OCCUPANCY_FOR(SURR) void fused_step_kernel(FusedArgs a) {
  const GridDev& g = a.grid;
  const TreeDev& tr = a.tree;
  const RateTables& rt = a.rates;
  const int lane = threadIdx.x;
  const int bx = a.tile_bits[0], by = a.tile_bits[1], bz = a.tile_bits[2];
  const int TX = 1 << bx, TY = 1 << by, TZ = 1 << bz;
  const int ntx = (g.nx + TX - 1) / TX, nty = (g.ny + TY - 1) / TY, ntz = (g.nz + TZ - 1) / TZ;
  const int64_t ntile = (int64_t)ntx * nty * ntz;
  // Tiles are handed out in Morton (Z-curve) order over the smallest
  // power-of-two cube of tiles that holds the mesh, whose places outside
  // the mesh are skipped: each XCD's queue is then an octant of that cube,
  // and the wavefronts running at one time on an XCD hold a compact block
  // of targets whose rays cross the same cells, so that more of their
  // gathers hit the XCD's L2 (1.08x on the 256^3 / 4096-source step).
  // A mesh whose cube would be more than 8 times its tile count keeps the
  // linear order.
  int mbits = 0;
  while ((1 << mbits) < max(ntx, max(nty, ntz))) mbits++;
  const int64_t ncube = (int64_t)1 << (3 * mbits);
  const bool morton = mbits <= 10 && ncube <= 8 * ntile;
  const int64_t nqueue_tiles = morton ? ncube : ntile;
  int q = blockIdx.x % kNumQueues, tried = 0;
  // per-lane counts (a lane's rays and crossings over the launch fit in 32
  // bits); wave-uniform visit counts
  unsigned int my_ray = 0, my_cross = 0, my_fb = 0, my_surr = 0;
  unsigned long long my_visit = 0, my_emit_iter = 0;
  // The lane's per-cell FP64 sums and its target's density and fractions
  // live in LDS between rays, so that the walk and the frequency loop have
  // the registers to themselves.
  __shared__ double acc_lds[6][64];
  // (the surrogate variant reads the target's fields from the mesh at each
  // ray instead, so that with the network's LDS it still fits 16 wavefronts
  // per CU)
  __shared__ float4 cell_lds[SURR ? 1 : 64];
  float* lds = nullptr;
  if constexpr (SURR) {
    __shared__ float surr_lds[kMlpLds];
    lds = surr_lds;
  }

  for (;;) {
    const int64_t t = next_tile(a.ctr, nqueue_tiles, &q, &tried);
    if (t < 0) break;
    int tx, ty, tz;
    if (morton) {
      tx = (int)compact3((uint32_t)t >> 2);
      ty = (int)compact3((uint32_t)t >> 1);
      tz = (int)compact3((uint32_t)t);
      if (tx >= ntx || ty >= nty || tz >= ntz) continue;   // outside the mesh
    } else {
      tz = (int)(t % ntz);
      ty = (int)((t / ntz) % nty);
      tx = (int)(t / ((int64_t)ntz * nty));
    }
    const int ix = tx * TX + ((lane >> (by + bz)) & (TX - 1));
    const int iy = ty * TY + ((lane >> bz) & (TY - 1));
    const int iz = tz * TZ + (lane & (TZ - 1));
    const bool valid = ix < g.nx && iy < g.ny && iz < g.nz;
    const int64_t cell = (int64_t)iz + (int64_t)g.nz * ((int64_t)iy + (int64_t)g.ny * ix);
    const Vec3f tgt{cell_centre(g.gxmin, ix + g.ox, g.dx), cell_centre(g.gymin, iy + g.oy, g.dy),
                    cell_centre(g.gzmin, iz + g.oz, g.dz)};
    {
      float4 f = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (valid)
        f = make_float4(*mesh_f(a.mesh, cell, kOffDens), *mesh_f(a.mesh, cell, kOffFHI),
                        *mesh_f(a.mesh, cell, kOffFHeI), *mesh_f(a.mesh, cell, kOffFHeII));
      if constexpr (!SURR) cell_lds[lane] = f;
#pragma unroll
      for (int k = 0; k < 6; k++) acc_lds[k][lane] = 0.0;
    }
    int p = valid ? 0 : tr.nnode;
    int seq = 0;
    for (;;) {
      const int u = wave_min(p);
      if (u >= tr.nnode) break;
      const float cx = tr.node_cm[4 * u + 0], cy = tr.node_cm[4 * u + 1],
                  cz = tr.node_cm[4 * u + 2], len2 = tr.node_cm[4 * u + 3];
      const NodeLink ln = tr.link[u];
      const bool here = (p == u);
      // setup_light_ray_to (setup_light_ray_tree.c:346-354), FP32
      const float ddx = tgt.x - cx, ddy = tgt.y - cy, ddz = tgt.z - cz;
      const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
      const bool accept = r2 * g.theta_sq > len2;
      const bool leaf = ln.n <= 1;
      const bool emit = here && (accept || ln.n == 1);
      if (here) p = (accept || leaf || ln.first < 0) ? ln.skip : ln.first;
      my_visit++;
      my_emit_iter += __ballot(emit) != 0 ? 1ull : 0ull;
      Vec3f src{0.0f, 0.0f, 0.0f}, st{0.0f, 0.0f, 0.0f};
      WalkOut w{};
      RayRate rr{};
      double res[6] = {0, 0, 0, 0, 0, 0};
      if (emit) {
        if (accept) {
          src = Vec3f{cx, cy, cz};
        } else {
          src = Vec3f{tr.src_pos[4 * ln.src + 0], tr.src_pos[4 * ln.src + 1],
                      tr.src_pos[4 * ln.src + 2]};
        }
        // a ray the reference refuses (calc_ray_segment.c:212) is counted
        // where it happens, on the path of sources outside the box
        segment_start(g, src, tgt, &st, &a.ctr->noverflow);
        walk_segment<KAHAN, CAPTURE, NARROW, POW2>(g, a.nf, st, tgt, &w);
        float4 f;
        if constexpr (SURR)
          f = make_float4(*mesh_f(a.mesh, cell, kOffDens), *mesh_f(a.mesh, cell, kOffFHI),
                          *mesh_f(a.mesh, cell, kOffFHeI), *mesh_f(a.mesh, cell, kOffFHeII));
        else
          f = cell_lds[lane];
        ray_rate_setup(g, rt, w.tau, src, tgt, f.x, f.y, f.z, f.w, &rr);
      }
      bool exact = emit;
      if constexpr (SURR) {
        // The network runs on the whole wavefront (MFMA) and serves the
        // visit only when every emitting lane can use it; otherwise every
        // lane takes the exact integral below.  Two 64x64 network
        // evaluations cost about as much as the exact 256-bin loop, so a
        // visit that ran both would cost twice as much.
        const double qn = a.surr.node_q[u];
        float xa[3] = {0.0f, 0.0f, 0.0f}, xb[3] = {0.0f, 0.0f, 0.0f};
        const bool use = emit && surr_inputs(a.surr, rt, rr, qn, xa, xb);
        const unsigned long long emitting = __ballot(emit);
        if (emitting != 0 && __ballot(use) == emitting) {
          float ya[kNout], yb[kNout];
          mlp_forward(a.surr.net, xa[0], xa[1], xa[2], ya, lds);
          mlp_forward(a.surr.net, xb[0], xb[1], xb[2], yb, lds);
          if (use) {
            surr_rates(rr, qn, ya, yb, res);
            exact = false;
            my_surr++;
          }
        }
      }
      if (exact) {
        if (fast_rate_ok(rt, rr)) {
          const float* sp = tr.spec + (int64_t)u * rt.ngrid;
          rate_loop_f32(rt, rr, ldc(tr.spec_scale + u), [&](int k) { return ldc(sp + k); }, res);
        } else {
          // the node spectrum as spec * spec_scale (within 6e-8 of the
          // double sums; this loop serves the rays FP32 cannot)
          const float* sp = tr.spec + (int64_t)u * rt.ngrid;
          const double sc = ldc(tr.spec_scale + u);
          rate_loop_f64(rt, rr, [&](int k) { return (double)ldc(sp + k) * sc; }, res);
          my_fb++;
        }
      }
      if (emit) {
#pragma unroll
        for (int k = 0; k < 6; k++) acc_lds[k][lane] += res[k];
        my_ray++;
        my_cross += (unsigned int)w.ncross;
        if (CAPTURE) {
          const unsigned long long slot = atomicAdd(&a.ctr->nrecord, 1ull);
          if (slot < (unsigned long long)a.rec_cap) {
            RayRecord& R = a.rec[slot];
            R.target = cell;
            R.seq = seq;
            R.kind = accept ? 0 : 1;
            R.id = accept ? u : ln.src;
            R.ncross = w.ncross;
            R.start[0] = st.x; R.start[1] = st.y; R.start[2] = st.z;
            R.src[0] = src.x; R.src[1] = src.y; R.src[2] = src.z;
            R.tau[0] = w.tau[0]; R.tau[1] = w.tau[1]; R.tau[2] = w.tau[2];
            R.surrogate = exact ? 0.0f : 1.0f;
            R.hash = w.hash;
            for (int k = 0; k < 6; k++) R.rate[k] = res[k];
          }
        }
        seq++;
      }
    }
    if (valid) {
      float* r = mesh_fw(a.mesh, cell, kOffRates);
#pragma unroll
      for (int k = 0; k < 6; k++) r[k] = (float)((double)r[k] + acc_lds[k][lane]);
    }
  }
  const unsigned long long sr = wave_sum((unsigned long long)my_ray),
                           sc = wave_sum((unsigned long long)my_cross),
                           sf = wave_sum((unsigned long long)my_fb),
                           ss = wave_sum((unsigned long long)my_surr);
  if (lane == 0) {
    atomicAdd(&a.ctr->nray, sr);
    atomicAdd(&a.ctr->ncross, sc);
    atomicAdd(&a.ctr->nfallback, sf);
    if (SURR) atomicAdd(&a.ctr->nsurrogate, ss);
    atomicAdd(&a.ctr->nvisit, my_visit);
    atomicAdd(&a.ctr->nemit_iter, my_emit_iter);
  }
}

// smooth_photoion_rate (smooth_photoion_rate.c:1058-1090) on one rank.
// The six prev_chem rates go to SoA planes, each plane is smoothed by the
// 27-point stencil (central weight 30, 1/56; ghost cells are the nearest
// interior cell), and the result goes back into prev_chem.
__global__ void gather_rates_kernel(const unsigned char* __restrict__ mesh, int64_t ncell,
                                    float* __restrict__ soa) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const float* r = mesh_f(mesh, c, kOffRates);
#pragma unroll
  for (int k = 0; k < 6; k++) soa[k * ncell + c] = r[k];
}

__global__ void smooth_rates_kernel(int nx, int ny, int nz, const float* __restrict__ in,
                                    float* __restrict__ out) {
  const int64_t ncell = (int64_t)nx * ny * nz;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 6 * ncell) return;
  const int k = (int)(i / ncell);
  const int64_t c = i - (int64_t)k * ncell;
  const int iz = (int)(c % nz), iy = (int)((c / nz) % ny), ix = (int)(c / ((int64_t)nz * ny));
  const float* p = in + (int64_t)k * ncell;
  // C evaluation order of the reference's sum: thirteen float additions,
  // then the double central term, then double additions
  float s = 0.0f;
  double d = 0.0;
  int term = 0;
#pragma unroll
  for (int dx = -1; dx <= 1; dx++) {
    const int jx = min(max(ix + dx, 0), nx - 1);
#pragma unroll
    for (int dy = -1; dy <= 1; dy++) {
      const int jy = min(max(iy + dy, 0), ny - 1);
#pragma unroll
      for (int dz = -1; dz <= 1; dz++) {
        const int jz = min(max(iz + dz, 0), nz - 1);
        const float v = p[(int64_t)jz + (int64_t)nz * ((int64_t)jy + (int64_t)ny * jx)];
        if (term == 0) s = v;
        else if (term < 13) s = s + v;
        else if (term == 13) d = (double)s + 30.0 * (double)v;
        else d = d + (double)v;
        term++;
      }
    }
  }
  out[i] = (float)((1.0 / (30.0 + 26.0)) * d);
}

__global__ void scatter_rates_kernel(unsigned char* __restrict__ mesh, int64_t ncell,
                                     const float* __restrict__ soa) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  float* r = mesh_fw(mesh, c, kOffRates);
#pragma unroll
  for (int k = 0; k < 6; k++) r[k] = soa[k * ncell + c];
}

hipError_t launch_smooth_rates(void* mesh, int nx, int ny, int nz, float* tmp_in, float* tmp_out,
                               hipStream_t s) {
  const int64_t ncell = (int64_t)nx * ny * nz;
  const unsigned nb = (unsigned)((ncell + 255) / 256);
  gather_rates_kernel<<<nb, 256, 0, s>>>(static_cast<unsigned char*>(mesh), ncell, tmp_in);
  smooth_rates_kernel<<<(unsigned)((6 * ncell + 255) / 256), 256, 0, s>>>(nx, ny, nz, tmp_in,
                                                                          tmp_out);
  scatter_rates_kernel<<<nb, 256, 0, s>>>(static_cast<unsigned char*>(mesh), ncell, tmp_out);
  return hipGetLastError();
}

// Plain device-to-device copy (16 B per thread per iteration, grid-stride):
// the bandwidth probe whose byte count calibrates the HBM counters.
__global__ void copy_probe_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                  int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

hipError_t launch_copy_probe(const float4* src, float4* dst, int64_t n, int nblock,
                             hipStream_t s) {
  copy_probe_kernel<<<nblock, 256, 0, s>>>(src, dst, n);
  return hipGetLastError();
}

__global__ void reset_counters_kernel(StepCounters* ctr) {
  const int i = threadIdx.x;
  if (i < kNumQueues) ctr->queue[i] = 0;
  if (i == 0) {
    ctr->nray = 0;
    ctr->ncross = 0;
    ctr->nfallback = 0;
    ctr->nrecord = 0;
    ctr->nsurrogate = 0;
    ctr->nvisit = 0;
    ctr->nemit_iter = 0;
    ctr->noverflow = 0;
  }
}

// ------------------------------------------------------------------------
// Launch wrappers (host side).

static inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_pack_fields(const void* mesh, int64_t ncell, double helium_fact, float4* nf,
                              hipStream_t s) {
  pack_fields_kernel<<<nblocks(ncell, 256), 256, 0, s>>>(
      static_cast<const unsigned char*>(mesh), ncell, helium_fact, nf);
  return hipGetLastError();
}

hipError_t launch_zero_rates(void* mesh, int64_t ncell, hipStream_t s) {
  zero_rates_kernel<<<nblocks(ncell, 256), 256, 0, s>>>(static_cast<unsigned char*>(mesh), ncell);
  return hipGetLastError();
}

hipError_t launch_add_rates(void* mesh, const float* rate, int64_t ncell, hipStream_t s) {
  add_rates_kernel<<<nblocks(ncell, 256), 256, 0, s>>>(static_cast<unsigned char*>(mesh), rate,
                                                        ncell);
  return hipGetLastError();
}

// Can the walk keep face positions instead of face counters?  Every cell
// size must be a power of two (then an integer face counter below 2^24
// times it is exact) and every mesh size below 2^23 (so that n + 1/2 is
// exact too).
static bool pow2_cells(const GridDev& g) {
  const float d[3] = {g.dx, g.dy, g.dz};
  const int n[3] = {g.nx, g.ny, g.nz};
  for (int k = 0; k < 3; k++) {
    int e = 0;
    if (!(d[k] >= FLT_MIN) || !(d[k] <= 1.0e30f) || std::frexp(d[k], &e) != 0.5f) return false;
    if (n[k] < 1 || n[k] >= (1 << 23)) return false;
  }
  return true;
}

template <typename Fields>
static void launch_walk(const GridDev& g, const Fields& f, void* seg, int64_t nseg, int kahan,
                        bool pow2, hipStream_t s) {
  if (pow2)
    walk_segments_kernel<true><<<nblocks(nseg, kWalkBlock), kWalkBlock, 0, s>>>(
        g, f, static_cast<unsigned char*>(seg), nseg, kahan);
  else
    walk_segments_kernel<false><<<nblocks(nseg, kWalkBlock), kWalkBlock, 0, s>>>(
        g, f, static_cast<unsigned char*>(seg), nseg, kahan);
}

hipError_t launch_walk_segments(const GridDev& g, const float4* nf, void* seg, int64_t nseg,
                                int kahan, int wide, hipStream_t s) {
  if (nseg <= 0) return hipSuccess;
  if (!wide && (int64_t)g.nx * g.ny * g.nz < kNarrowCells)
    launch_walk(g, PackedFields<true>{nf, (uint32_t)(g.nx * g.nyz * 16)}, seg, nseg, kahan,
                pow2_cells(g), s);
  else
    launch_walk(g, PackedFields<false>{nf, 0u}, seg, nseg, kahan, false, s);
  return hipGetLastError();
}

hipError_t launch_walk_segments_mesh(const GridDev& g, const void* mesh, double helium_fact,
                                     void* seg, int64_t nseg, int kahan, hipStream_t s) {
  if (nseg <= 0) return hipSuccess;
  launch_walk(g, MeshFields{static_cast<const unsigned char*>(mesh), helium_fact}, seg, nseg,
              kahan, pow2_cells(g), s);
  return hipGetLastError();
}

hipError_t launch_target_rates(const GridDev& g, const RateTables& rt, void* mesh,
                              const CallTarget* tg, int64_t ntarget, const CallRay* rays,
                              const float* spec32, const double* spec64, const double* spec_scale,
                              hipStream_t s) {
  if (ntarget <= 0) return hipSuccess;
  target_rates_kernel<<<nblocks(ntarget, 64), 64, 0, s>>>(
      g, rt, static_cast<unsigned char*>(mesh), tg, ntarget, rays, spec32, spec64, spec_scale);
  return hipGetLastError();
}

hipError_t launch_get_rates(const void* mesh, float* rate, int64_t ncell, hipStream_t s) {
  if (ncell <= 0) return hipSuccess;
  get_rates_kernel<<<nblocks(ncell, 256), 256, 0, s>>>(static_cast<const unsigned char*>(mesh),
                                                      ncell, rate);
  return hipGetLastError();
}

hipError_t launch_set_rates(void* mesh, const float* rate, int64_t ncell, hipStream_t s) {
  if (ncell <= 0) return hipSuccess;
  set_rates_kernel<<<nblocks(ncell, 256), 256, 0, s>>>(static_cast<unsigned char*>(mesh), rate,
                                                      ncell);
  return hipGetLastError();
}

hipError_t launch_reset_counters(StepCounters* ctr, hipStream_t s) {
  reset_counters_kernel<<<1, 64, 0, s>>>(ctr);
  return hipGetLastError();
}

template <bool SURR, bool NARROW, bool POW2>
static void launch_fused(const FusedArgs& a, int kahan, int capture, int nblock, hipStream_t s) {
  if (capture) {
    if (kahan)
      fused_step_kernel<true, true, SURR, NARROW, POW2><<<nblock, 64, 0, s>>>(a);
    else
      fused_step_kernel<false, true, SURR, NARROW, POW2><<<nblock, 64, 0, s>>>(a);
  } else {
    if (kahan)
      fused_step_kernel<true, false, SURR, NARROW, POW2><<<nblock, 64, 0, s>>>(a);
    else
      fused_step_kernel<false, false, SURR, NARROW, POW2><<<nblock, 64, 0, s>>>(a);
  }
}

// Cell-field addressing and face arithmetic: 32-bit offsets on meshes of
// fewer than 2^28 cells, with face positions where the cell sizes allow;
// otherwise (and always when wide is set) 64-bit addresses and face counters.
template <bool SURR>
static void launch_fused_form(const FusedArgs& a, int kahan, int capture, int wide, int nblock,
                              hipStream_t s) {
  const bool narrow = !wide && (int64_t)a.grid.nx * a.grid.ny * a.grid.nz < kNarrowCells;
  if (narrow && pow2_cells(a.grid))
    launch_fused<SURR, true, true>(a, kahan, capture, nblock, s);
  else if (narrow)
    launch_fused<SURR, true, false>(a, kahan, capture, nblock, s);
  else
    launch_fused<SURR, false, false>(a, kahan, capture, nblock, s);
}

hipError_t launch_fused_step(const FusedArgs& a, int kahan, int capture, int surr, int wide,
                             int nblock, hipStream_t s) {
  if (surr)
    launch_fused_form<true>(a, kahan, capture, wide, nblock, s);
  else
    launch_fused_form<false>(a, kahan, capture, wide, nblock, s);
  return hipGetLastError();
}

}  // namespace argot
