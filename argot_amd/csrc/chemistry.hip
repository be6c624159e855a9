// HIP kernels of the chemistry step for gfx950 (MI355X), one thread per cell.
//
//   chem_step     advance_reaction_and_heatcool_kernel (chemistry.cu:390-433):
//                 prev_chem/prev_uene -> chem/uene over dtime, maxima of the
//                 relative changes; one cell per thread, or persistent with a
//                 cell queue (chem_step_queue)
//   chem_dtime    calc_dtime_gpu_kernel (chemistry.cu:246-269): minimum safe
//                 step over cells
//   chem_update   update_chemistry_kernel (chemistry.cu:224-234): prev := new,
//                 total energy
//
// The cells are AoS struct fluid_mesh records (156 B).  The step is bound by
// double-precision arithmetic (transcendentals in the rate fits and the
// bisection), so each thread reads its record's fields directly.
#include <hip/hip_runtime.h>

#include <algorithm>

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

// One cell per thread, all of its sub-steps in turn.
__global__ void __launch_bounds__(kBlock) chem_step_kernel(unsigned char* __restrict__ mesh,
                                                           int64_t ncell, ChemParams p,
                                                           float dtime,
                                                           ChemCounters* __restrict__ ctr) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  unsigned int dchem = 0u, duene = 0u;
  unsigned long long nsub = 0, niter = 0, nstop = 0;
  if (c < ncell) {
    unsigned char* m = mesh + c * kStride;
    chem::CellRun r;
    r.dens = ld(m, kDens);
    r.mesh_uene = ld(m, kUene);
    r.uene = ld(m, kPrevUene);
    r.c = ld_state(m, kPrevChem);
    st(m, kDuene, 0.0f);
    chem::cell_begin(r, p, dtime);
    while (chem::cell_more(r)) chem::cell_substep(r, p);
    const chem::State old = ld_state(m, kChem);
    dchem = __float_as_uint(chem::chem_change(r.c, old));
    duene = __float_as_uint(chem::uene_change(r.uene, r.mesh_uene));
    st(m, kUene, r.uene);
    st_state(m, kChem, r.c);
    nsub = (unsigned long long)r.nsub;
    niter = (unsigned long long)r.niter;
    nstop = chem::cell_stopped(r) ? 1ull : 0ull;
  }
  wave_reduce_to<true>(dchem, &ctr->diff_chem);
  wave_reduce_to<true>(duene, &ctr->diff_uene);
  wave_sum_to(nsub, &ctr->nsub);
  wave_sum_to(niter, &ctr->niter);
  wave_sum_to(nstop, &ctr->nstopped);
}

// Persistent form: a lane whose cell is done takes the next one from a
// queue (one atomic per wavefront refill), so a wavefront stays full while
// a few of its cells need many sub-steps.  Each cell's arithmetic is the
// same as in chem_step_kernel; only the interleaving of cells differs.
__global__ void __launch_bounds__(kBlock) chem_step_queue_kernel(unsigned char* __restrict__ mesh,
                                                                 int64_t ncell, ChemParams p,
                                                                 float dtime, int min_refill,
                                                                 ChemCounters* __restrict__ ctr) {
  const int lane = threadIdx.x & 63;
  unsigned int dchem = 0u, duene = 0u;
  unsigned long long nsub = 0, niter = 0, nstop = 0;
  chem::CellRun r;
  int64_t cell = -1;       // the lane's current cell, -1: needs one
  bool drained = false;    // the queue had no cell left for this lane
  for (;;) {
    // Free lanes take new cells once min_refill of them are free, or when
    // no lane has a cell left: the set-up of a cell then runs for several
    // lanes at once instead of stalling the wavefront for one.
    const bool need = cell < 0 && !drained;
    const unsigned long long want = __ballot(need);
    if (want && (__popcll(want) >= (unsigned)min_refill || __all(cell < 0))) {
      const int leader = __builtin_ctzll(want);
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(&ctr->next, (unsigned long long)__popcll(want));
      base = __shfl(base, leader, 64);
      if (need) {
        const unsigned long long idx =
            base + (unsigned long long)__popcll(want & ((1ull << lane) - 1ull));
        if (idx < (unsigned long long)ncell) {
          cell = (int64_t)idx;
          unsigned char* m = mesh + cell * kStride;
          r.dens = ld(m, kDens);
          r.mesh_uene = ld(m, kUene);
          r.uene = ld(m, kPrevUene);
          r.c = ld_state(m, kPrevChem);
          st(m, kDuene, 0.0f);
          chem::cell_begin(r, p, dtime);
        } else {
          drained = true;
        }
      }
    }
    if (__all(cell < 0)) break;   // every lane drained and idle
    if (cell >= 0) {
      if (chem::cell_more(r)) {
        chem::cell_substep(r, p);
      } else {
        unsigned char* m = mesh + cell * kStride;
        const chem::State old = ld_state(m, kChem);
        const unsigned int a = __float_as_uint(chem::chem_change(r.c, old));
        const unsigned int b = __float_as_uint(chem::uene_change(r.uene, r.mesh_uene));
        dchem = a > dchem ? a : dchem;
        duene = b > duene ? b : duene;
        st(m, kUene, r.uene);
        st_state(m, kChem, r.c);
        nsub += (unsigned long long)r.nsub;
        niter += (unsigned long long)r.niter;
        nstop += chem::cell_stopped(r) ? 1ull : 0ull;
        cell = -1;
      }
    }
  }
  wave_reduce_to<true>(dchem, &ctr->diff_chem);
  wave_reduce_to<true>(duene, &ctr->diff_uene);
  wave_sum_to(nsub, &ctr->nsub);
  wave_sum_to(niter, &ctr->niter);
  wave_sum_to(nstop, &ctr->nstopped);
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
    ctr->next = 0ull;
    ctr->nstopped = 0ull;
  }
}

unsigned int nblocks(int64_t ncell) { return (unsigned int)((ncell + kBlock - 1) / kBlock); }

}  // namespace

hipError_t launch_chem_step(void* mesh, int64_t ncell, const ChemParams& p, float dtime,
                            ChemCounters* ctr, int queue_blocks, int min_refill, hipStream_t s) {
  chem_reset_kernel<<<1, 64, 0, s>>>(ctr);
  if (ncell <= 0) return hipGetLastError();
  if (queue_blocks > 0) {
    const unsigned int nb = (unsigned int)std::min<int64_t>(queue_blocks, nblocks(ncell));
    chem_step_queue_kernel<<<nb, kBlock, 0, s>>>((unsigned char*)mesh, ncell, p, dtime,
                                                 std::max(min_refill, 1), ctr);
  } else {
    chem_step_kernel<<<nblocks(ncell), kBlock, 0, s>>>((unsigned char*)mesh, ncell, p, dtime, ctr);
  }
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
