// HIP kernels of the chemistry step for gfx950 (MI355X), one thread per cell.
//
//   chem_step     advance_reaction_and_heatcool_kernel (chemistry.cu:390-433):
//                 prev_chem/prev_uene -> chem/uene over dtime, maxima of the
//                 relative changes
//   chem_dtime    calc_dtime_gpu_kernel (chemistry.cu:246-269): minimum safe
//                 step over cells
//   chem_update   update_chemistry_kernel (chemistry.cu:224-234): prev := new,
//                 total energy
//
// The cells are AoS struct fluid_mesh records (156 B).  The step is bound by
// double-precision arithmetic (transcendentals in the rate fits and the
// bisection), so each thread reads its record's fields directly.
#include <hip/hip_runtime.h>

#include "argot_internal.h"
#include "chem_device.h"
#include "kernels.h"

namespace argot {

namespace {

constexpr int kStride = 156;        // sizeof(struct fluid_mesh)
constexpr int kDens = 0, kEneg = 4, kMomX = 8, kMomY = 12, kMomZ = 16;
constexpr int kUene = 20, kDuene = 24;
constexpr int kChem = 44, kPrevChem = 96, kPrevUene = 148;
constexpr int kBlock = 256;

__device__ __forceinline__ float ld(const unsigned char* m, int off) {
  return *reinterpret_cast<const float*>(m + off);
}
__device__ __forceinline__ void st(unsigned char* m, int off, float v) {
  *reinterpret_cast<float*>(m + off) = v;
}
__device__ __forceinline__ chem::State ld_state(const unsigned char* m, int off) {
  chem::State s;
  float* d = &s.wmol;
#pragma unroll
  for (int k = 0; k < 13; k++) d[k] = ld(m, off + 4 * k);
  return s;
}
__device__ __forceinline__ void st_state(unsigned char* m, int off, const chem::State& s) {
  const float* d = &s.wmol;
#pragma unroll
  for (int k = 0; k < 13; k++) st(m, off + 4 * k, d[k]);
}

// Wave-wide maximum (mx) or minimum (!mx) of an unsigned value, then one
// atomic per wave.
template <bool MX>
__device__ __forceinline__ void wave_reduce_to(unsigned int v, unsigned int* dst) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned int w = __shfl_xor(v, o, 64);
    v = MX ? (w > v ? w : v) : (w < v ? w : v);
  }
  if ((threadIdx.x & 63) == 0) {
    if (MX)
      atomicMax(dst, v);
    else
      atomicMin(dst, v);
  }
}

__device__ __forceinline__ void wave_sum_to(unsigned long long v, unsigned long long* dst) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(dst, v);
}

__global__ void __launch_bounds__(kBlock) chem_step_kernel(unsigned char* __restrict__ mesh,
                                                           int64_t ncell, ChemParams p,
                                                           float dtime,
                                                           ChemCounters* __restrict__ ctr) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  unsigned int dchem = 0u, duene = 0u;
  unsigned long long nsub = 0, niter = 0;
  if (c < ncell) {
    unsigned char* m = mesh + c * kStride;
    const float dens = ld(m, kDens);
    const float mesh_uene = ld(m, kUene);
    float uene = ld(m, kPrevUene);
    chem::State s = ld_state(m, kPrevChem);
    st(m, kDuene, 0.0f);
    const chem::StepOut o = chem::advance_reaction_and_heatcool(dens, mesh_uene, &uene, &s, p, dtime);
    const chem::State old = ld_state(m, kChem);
    dchem = __float_as_uint(chem::chem_change(s, old));
    duene = __float_as_uint(chem::uene_change(uene, mesh_uene));
    st(m, kUene, uene);
    st_state(m, kChem, s);
    nsub = (unsigned long long)o.nsub;
    niter = (unsigned long long)o.niter;
  }
  wave_reduce_to<true>(dchem, &ctr->diff_chem);
  wave_reduce_to<true>(duene, &ctr->diff_uene);
  wave_sum_to(nsub, &ctr->nsub);
  wave_sum_to(niter, &ctr->niter);
}

__global__ void __launch_bounds__(kBlock) chem_dtime_kernel(const unsigned char* __restrict__ mesh,
                                                            int64_t ncell, ChemParams p,
                                                            ChemCounters* __restrict__ ctr) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  unsigned int dt = 0x7f7fffffu;   // FLT_MAX
  if (c < ncell) {
    const unsigned char* m = mesh + c * kStride;
    const chem::State s = ld_state(m, kPrevChem);
    const chem::Dens n = chem::number_densities(ld(m, kDens), s, p);
    const float T = chem::temperature(ld(m, kUene), s, p);
    dt = __float_as_uint(chem::safe_dtime(n, s, chem::rate_coefficients(T)));
  }
  wave_reduce_to<false>(dt, &ctr->dt_min);
}

__global__ void chem_update_kernel(unsigned char* __restrict__ mesh, int64_t ncell) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  unsigned char* m = mesh + c * kStride;
  st_state(m, kPrevChem, ld_state(m, kChem));
  const float uene = ld(m, kUene), dens = ld(m, kDens);
  st(m, kPrevUene, uene);
  const float mx = ld(m, kMomX), my = ld(m, kMomY), mz = ld(m, kMomZ);
  st(m, kEneg,
     (float)((double)(uene * dens) + 0.5 * (double)(mx * mx + my * my + mz * mz) / (double)dens));
}

__global__ void chem_reset_kernel(ChemCounters* ctr) {
  if (threadIdx.x == 0) {
    ctr->diff_chem = 0u;
    ctr->diff_uene = 0u;
    ctr->dt_min = 0x7f7fffffu;
    ctr->pad = 0u;
    ctr->nsub = 0ull;
    ctr->niter = 0ull;
  }
}

unsigned int nblocks(int64_t ncell) { return (unsigned int)((ncell + kBlock - 1) / kBlock); }

}  // namespace

hipError_t launch_chem_step(void* mesh, int64_t ncell, const ChemParams& p, float dtime,
                            ChemCounters* ctr, hipStream_t s) {
  chem_reset_kernel<<<1, 64, 0, s>>>(ctr);
  if (ncell > 0)
    chem_step_kernel<<<nblocks(ncell), kBlock, 0, s>>>((unsigned char*)mesh, ncell, p, dtime, ctr);
  return hipGetLastError();
}

hipError_t launch_chem_dtime(const void* mesh, int64_t ncell, const ChemParams& p,
                             ChemCounters* ctr, hipStream_t s) {
  chem_reset_kernel<<<1, 64, 0, s>>>(ctr);
  if (ncell > 0)
    chem_dtime_kernel<<<nblocks(ncell), kBlock, 0, s>>>((const unsigned char*)mesh, ncell, p, ctr);
  return hipGetLastError();
}

hipError_t launch_chem_update(void* mesh, int64_t ncell, hipStream_t s) {
  if (ncell > 0)
    chem_update_kernel<<<nblocks(ncell), kBlock, 0, s>>>((unsigned char*)mesh, ncell);
  return hipGetLastError();
}

}  // namespace argot
