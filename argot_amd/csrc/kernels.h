// Launch interface of kernels.hip (host side of libargot_hip.so).
#pragma once
#include <hip/hip_runtime.h>

#include "argot_internal.h"
#include "surrogate.h"

namespace argot {

struct FusedArgs {
  GridDev grid;
  TreeDev tree;
  RateTables rates;
  const float4* nf;            // per-cell (nHI, nHeI, nHeII, 0)
  unsigned char* mesh;         // AoS fluid_mesh (reads fractions, adds rates)
  StepCounters* ctr;
  RayRecord* rec;              // capture mode only
  int64_t rec_cap;
  SurrDev surr;                // rate surrogate (surrogate launches only)
  int tile_bits[3];            // target tile of a wavefront: 2^bx x 2^by x 2^bz = 64 cells
};

hipError_t launch_pack_fields(const void* mesh, int64_t ncell, double helium_fact, float4* nf,
                              hipStream_t s);
hipError_t launch_zero_rates(void* mesh, int64_t ncell, hipStream_t s);
hipError_t launch_add_rates(void* mesh, const float* rate, int64_t ncell, hipStream_t s);
// rate[6c + k] := prev_chem rate k of cell c, and back
hipError_t launch_get_rates(const void* mesh, float* rate, int64_t ncell, hipStream_t s);
hipError_t launch_set_rates(void* mesh, const float* rate, int64_t ncell, hipStream_t s);
// wide: address the cell fields in 64 bits even when the mesh has fewer
// than 2^28 cells (meshes with more always are).
hipError_t launch_walk_segments(const GridDev& g, const float4* nf, void* seg, int64_t nseg,
                                int kahan, int wide, hipStream_t s);
// The same walk with the densities formed from the AoS mesh records at each
// crossing (no packed-field buffer; identical optical depths).
hipError_t launch_walk_segments_mesh(const GridDev& g, const void* mesh, double helium_fact,
                                     void* seg, int64_t nseg, int kahan, hipStream_t s);
// calc_photoion_rate_cuda: ntarget targets, each with its consecutive rays
// (rays[first .. first + count)), rates added to the mesh's prev_chem.
// spec32/spec64/spec_scale: the distinct spectra ([nspec * ngrid] FP32
// scaled by 1/spec_scale, FP64 as given, and the per-spectrum maximum).
hipError_t launch_target_rates(const GridDev& g, const RateTables& rt, void* mesh,
                              const CallTarget* tg, int64_t ntarget, const CallRay* rays,
                              const float* spec32, const double* spec64, const double* spec_scale,
                              hipStream_t s);
hipError_t launch_reset_counters(StepCounters* ctr, hipStream_t s);
// smooth_photoion_rate on one rank; tmp_in/tmp_out: 6*ncell floats each.
// On return (stream order) tmp_out holds the smoothed rates as six planes.
hipError_t launch_smooth_rates(void* mesh, int nx, int ny, int nz, float* tmp_in, float* tmp_out,
                               hipStream_t s);
hipError_t launch_copy_probe(const float4* src, float4* dst, int64_t n, int nblock,
                             hipStream_t s);
hipError_t launch_fused_step(const FusedArgs& a, int kahan, int capture, int surr, int wide,
                             int nblock, hipStream_t s);

// Chemistry (chemistry.hip).  Each launch first resets *ctr on the stream.
// queue_blocks > 0: the persistent cell-queue kernel on that many blocks;
// 0: one cell per thread.
hipError_t launch_chem_step(void* mesh, int64_t ncell, const ChemParams& p, float dtime,
                            ChemCounters* ctr, int queue_blocks, int min_refill, hipStream_t s);
hipError_t launch_chem_dtime(const void* mesh, int64_t ncell, const ChemParams& p,
                             ChemCounters* ctr, hipStream_t s);
hipError_t launch_chem_update(void* mesh, int64_t ncell, hipStream_t s);

}  // namespace argot
