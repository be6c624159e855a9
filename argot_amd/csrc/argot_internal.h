// Internal types shared by the host runtime and the HIP kernels of
// libargot_hip.so.  Not part of the C-ABI (include/argot_hip.h is).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ARGOT_HD __host__ __device__
#else
#define ARGOT_HD
struct float4_host {
  float x, y, z, w;
};
#endif

namespace argot {

constexpr int kMaxSegments = 5;   // NSEG_PER_RAY (run_param.h:60)
constexpr int kNgridMax = 256;    // NGRID_NU (run_param.h:78)
constexpr int kTileBits = 6;      // 64 target cells (one per lane) per wavefront tile
constexpr int kNumQueues = 8;     // one tile queue per XCD

// Mesh geometry and physical units as the kernels use them.
struct GridDev {
  int nx, ny, nz;                            // local mesh
  int64_t nyz;
  float xmin, ymin, zmin, xmax, ymax, zmax;  // local box (walk, calc_optical_depth.c:76-143)
  float gxmin, gymin, gzmin;                 // global origin (target centres)
  int ox, oy, oz;                            // global index of local cell 0
  float dx, dy, dz;
  double lunit, denstonh, helium_fact;
  float theta_sq;   // SQR(theta_crit) in FP32 (setup_light_ray_tree.c:577)
};

// Frequency tables (device pointers).  sig/hsig are [3][ngrid] SoA:
// sig[s][k] = sigma_s(nu_k), hsig[s][k] = sigma_s(nu_k) (nu_k - nu_s) for
// nu_k > nu_s else 0, with nu_k the FP32 grid value as the reference uses.
struct RateTables {
  int ngrid;
  int k_he1;           // bins below k_he1 have sigma_HeI = sigma_HeII = 0
  int k_he2;           // bins below k_he2 have sigma_HeII = 0
  const float* nu;     // [ngrid]
  const float* cs;     // [3*ngrid], bin-major as run_param.csect
  const float* sig;    // [3*ngrid], species-major
  const float* hsig;   // [3*ngrid], species-major
  const float* dsig;   // [3*ngrid] sigma_s(k) - sigma_s(last bin), species-major
  double cs_max[3];    // max over bins of sigma_s
  double cs_last[3];   // sigma_s at the last bin
};

// Source tree in depth-first pre-order (children in octant order), as
// setup_light_ray_to (setup_light_ray_tree.c:321-395) visits it.
struct NodeLink {
  int skip;       // next node after this subtree (== nnode at the end)
  int n;          // number of sources below
  int first;      // first child (== self+1) or -1 when there is none
  int src;        // for n == 1: the source (original index), else -1
};

struct TreeDev {
  int nnode;
  const float* node_cm;      // [4*nnode]: cm x, y, z and SQR(length)
  const NodeLink* link;      // [nnode]
  const float* spec;         // [nnode*ngrid] spectrum / spec_scale, FP32
  const double* spec_scale;  // [nnode] max over bins of the spectrum
  const float* src_pos;      // [4*nsrc]
};

// One emitted ray, written only in capture mode (tests).
struct RayRecord {
  int64_t target;       // linear target index
  int32_t seq;          // position in the target's ray list
  int32_t kind;         // 0: tree node, 1: source
  int32_t id;           // node (pre-order) or source index
  int32_t ncross;       // crossings incl. the end cell
  float start[3];       // segment start (global coordinates)
  float src[3];         // ray source position
  float tau[3];         // FP32 optical depths (code units)
  float surrogate;      // 1: rates from the network, 0: exact integral
  uint64_t hash;        // FNV-1a of the visited cell sequence
  double rate[6];       // this ray's rates (heats times hplanck*nuL)
};

// calc_photoion_rate_cuda inputs after the host's regrouping: one record
// per ray (32 B) and one per target cell (32 B).
struct CallRay {
  float src[3];      // ray source position
  float tau[3];      // optical depths (code units)
  int32_t spec;      // index of the ray's spectrum among the call's distinct ones
  int32_t pad;
};
struct CallTarget {
  int64_t first;     // its first ray
  int64_t cell;      // local linear cell index
  int32_t count;     // its rays
  int32_t g[3];      // global cell indices (target centre)
};

struct StepCounters {
  unsigned long long nray;
  unsigned long long ncross;
  unsigned long long nfallback;
  unsigned long long nrecord;
  unsigned long long nsurrogate;
  unsigned long long nvisit;       // tree nodes visited, summed over wavefronts
  unsigned long long nemit_iter;   // of those, visits where some lane emitted a ray
  unsigned long long noverflow;    // rays that would need NSEG_PER_RAY or more segments
  int queue[kNumQueues];
};

// Run-time parameters of the chemistry (from struct run_param, plus the
// helium number fraction).
struct ChemParams {
  double lunit, munit, tunit, uenetok, denstonh, helium_fact;
  float anow, znow;
};

// Results of one chemistry launch: maxima over cells as float bits (the
// values are non-negative, so unsigned order is float order), the minimum
// safe step, and work counts.
struct ChemCounters {
  unsigned int diff_chem;
  unsigned int diff_uene;
  unsigned int dt_min;
  unsigned int pad;
  unsigned long long nsub;
  unsigned long long niter;
  unsigned long long next;   // cell queue of the persistent chemistry kernel
  unsigned long long nstopped;   // cells ended by the NaN or sub-step guard
};

}  // namespace argot
