// Device-side building blocks of the point-source radiation step (gfx950).
//
// Everything here reproduces the reference's arithmetic operation by
// operation: the file is compiled with -ffp-contract=off and without
// fast-math, so a*b+c stays two roundings, float division and sqrt are
// IEEE, and every double promotion of the C source is written out.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "argot_internal.h"

namespace argot {

constexpr double kWalkTiny = 1.0e-31;   // calc_optical_depth.c:15
constexpr double kSegTiny = 1.0e-30;    // calc_ray_segment.c:12
constexpr double kRateTiny = 1.0e-31;   // calc_photoion_rate.c:16
constexpr double kPi = 3.1415926535897932384626433;  // constants.h:24
constexpr double kHnuL = 6.626e-27 * 3.28e15;          // hplanck*nuL

// Load through the constant address space: with a wave-uniform address the
// compiler emits a scalar load (s_load), keeping tables out of VGPRs and off
// the vector memory path.  Only for memory no kernel writes while it runs.
template <typename T>
__device__ __forceinline__ T ldc(const T* p) {
  return *(const __attribute__((address_space(4))) T*)(p);
}

// Quotient a/b rounded to float exactly as IEEE float division does, from
// the double reciprocal rb = 1.0/(double)b (rounded to double).  The
// double product is within 2^-52 (relative) of a/b, while a quotient of two
// 24-bit floats is never closer than 2^-49 (relative) to a rounding
// boundary of float, so rounding the product to float gives RN(a/b).
__device__ __forceinline__ float div_by(float a, double rb) {
  return (float)((double)a * rb);
}

// Target cell centre, setup_light_ray_tree.c:341-343 / calc_ray_segment.c:48:
// (float)i + 0.5 is double, the sum is rounded to float once.
ARGOT_HD inline float cell_centre(float lo, int i, float h) {
  return (float)((double)lo + ((double)(float)i + 0.5) * (double)h);
}

struct Vec3f {
  float x, y, z;
};

// calc_ray_segment.c:15-213 for one rank: start point of the segment whose
// optical depth the ray keeps (a source outside the box enters at a face;
// the domain loop can split that entry into several local segments, of
// which accum_optical_depth.c:93-101 keeps the last).  Returns the number of
// segments the reference forms, counted up to kMaxSegments: the reference
// asserts num_segment < NSEG_PER_RAY (calc_ray_segment.c:212), so a return
// of kMaxSegments means the ray is refused (its start point is then the
// one after kMaxSegments - 1 domain steps, not the reference's); on the
// device such a ray also adds one to *refused when that is given.  Host and
// device compile the same IEEE operations, so both give the same bits.
ARGOT_HD inline int segment_start(const GridDev& g, Vec3f src, Vec3f tgt, Vec3f* start,
                                  unsigned long long* refused = nullptr) {
  const float dxd = (g.xmax - g.xmin) / 1.0f;
  const float dyd = (g.ymax - g.ymin) / 1.0f;
  const float dzd = (g.zmax - g.zmin) / 1.0f;
  const int ixt = (int)((tgt.x - g.xmin) / dxd);
  const int iyt = (int)((tgt.y - g.ymin) / dyd);
  const int izt = (int)((tgt.z - g.zmin) / dzd);
  const bool inside = !(src.x < g.xmin || src.x > g.xmax || src.y < g.ymin || src.y > g.ymax ||
                        src.z < g.zmin || src.z > g.zmax);
  if (inside) {
    const int ixs = (int)((src.x - g.xmin) / dxd);
    const int iys = (int)((src.y - g.ymin) / dyd);
    const int izs = (int)((src.z - g.zmin) / dzd);
    if (ixs == ixt && iys == iyt && izs == izt) {
      *start = src;
      return 1;
    }
  }
  float dx = tgt.x - src.x, dy = tgt.y - src.y, dz = tgt.z - src.z;
  float dl = (float)__builtin_sqrt((double)(dx * dx + dy * dy + dz * dz) + kSegTiny);
  float cos_t = dz / dl;
  float sin_t = (float)__builtin_sqrt(1.0 - (double)(cos_t * cos_t));
  float dxy = dl * sin_t;
  float cos_p = (float)((double)dx / ((double)dxy + kSegTiny));
  float sin_p = (float)((double)dy / ((double)dxy + kSegTiny));
  float xo = cos_p * sin_t;
  if ((double)fabsf(xo) < kSegTiny) xo = (float)((double)xo >= 0.0 ? kSegTiny : -kSegTiny);
  float yo = sin_p * sin_t;
  if (fabs((double)yo) < kSegTiny) yo = (float)((double)yo >= 0.0 ? kSegTiny : -kSegTiny);
  float zo = cos_t;
  if (fabs((double)zo) < kSegTiny) zo = (float)((double)zo >= 0.0 ? kSegTiny : -kSegTiny);
  const int xs = (double)xo > 0.0, ys = (double)yo > 0.0, zs = (double)zo > 0.0;
  const int xn = xs ? 1 : -1, yn = ys ? 1 : -1, zn = zs ? 1 : -1;
  float out = 0.0f;
  if (src.x < g.xmin) out = (g.xmin - src.x) / xo;
  if (src.x > g.xmax) out = (g.xmax - src.x) / xo;
  if (src.y < g.ymin) out = (g.ymin - src.y) / yo;
  if (src.y > g.ymax) out = (g.ymax - src.y) / yo;
  if (src.z < g.zmin) out = (g.zmin - src.z) / zo;
  if (src.z > g.zmax) out = (g.zmax - src.z) / zo;
  float sx, sy, sz;
  if ((double)out != 0.0) {
    sx = src.x + out * xo;
    sy = src.y + out * yo;
    sz = src.z + out * zo;
  } else {
    sx = src.x;
    sy = src.y;
    sz = src.z;
  }
  int ixs = (int)((sx - g.xmin) / dxd);
  int iys = (int)((sy - g.ymin) / dyd);
  int izs = (int)((sz - g.zmin) / dzd);
  int nseg = 1;
  while ((ixs != ixt || iys != iyt || izs != izt) && nseg < kMaxSegments) {
    float ddx = g.xmin + dxd * (float)(xs + ixs) - sx;
    float ddy = g.ymin + dyd * (float)(ys + iys) - sy;
    float ddz = g.zmin + dzd * (float)(zs + izs) - sz;
    float rx = ddx / xo, ry = ddy / yo, rz = ddz / zo;
    float rmin = fminf(rx, fminf(ry, rz));
    if (rmin == rx) ixs += xn;
    if (rmin == ry) iys += yn;
    if (rmin == rz) izs += zn;
    sx = sx + rmin * xo;
    sy = sy + rmin * yo;
    sz = sz + rmin * zo;
    nseg++;
  }
#if defined(__HIP_DEVICE_COMPILE__)
  if (refused && nseg >= kMaxSegments) atomicAdd(refused, 1ull);
#endif
  start->x = sx;
  start->y = sy;
  start->z = sz;
  return nseg;
}

// Per-ray state of the DDA walk, calc_optical_depth.c:76-143.  Along each
// axis the walk keeps, instead of the cell index i, the face counter
// f = sign * ((float)i + step) (step = 1 when moving up, else 0; sign = +-1
// the direction), an integer-valued float that grows by 1 at every crossing
// of that axis: sign * delta * f == delta * ((float)i + step) exactly, and
// the cell is inside the mesh on that axis while f < lim (lim = n + 1 when
// moving up, 1 when moving down: the walk never moves back).  The linear
// cell index moves by the axis stride at each crossing.
//
// POW2 (cell sizes that are powers of two): f holds the face position
// sign * delta * f itself, moved by sign * delta at each crossing, which is
// exact (an integer below 2^24 times a power of two), so the face distances
// are bit for bit those of the general form without its multiply.  The cell
// is inside on an axis while (f < l) != neg, with l = (n + 1/2) delta when
// moving up and -delta/2 when moving down (f is a multiple of delta, so the
// comparison never meets equality).
struct Dda {
  float fx, fy, fz;      // face counters (POW2: face positions)
  float sdx, sdy, sdz;   // sign * delta
  float lx, ly, lz;      // inside while f < l (POW2: while (f < l) != neg)
  bool negx, negy, negz; // POW2: moving down
  uint32_t stx, sty, stz;  // strides of the linear index times S, modulo 2^32
  uint32_t idx;          // linear index of the current cell times S
  uint32_t idxe;         // linear index of the end cell times S
  float xo, yo, zo;      // direction cosines (with the TINY guards)
  double rxo, ryo, rzo;  // 1/xo etc. in double (exact-division helper)
  float xc, yc, zc;      // position in local coordinates
};

// Linear index iz + nz (iy + ny ix) of a cell inside the mesh, from two
// 24-bit multiply-adds (mesh sizes and indices are below 2^24).  The inner
// one is written out: left to itself the compiler folds it into a
// quarter-rate 64-bit multiply-add.  ny and nz are kernel arguments, so
// they sit in scalar registers.
__device__ __forceinline__ uint32_t cell_index(const GridDev& g, int ix, int iy, int iz) {
  uint32_t row;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(row) : "v"(ix), "s"(g.ny), "v"(iy));
  return __umul24(row, (uint32_t)g.nz) + (uint32_t)iz;
}

// S scales the linear indices the walk carries (16: byte offsets into the
// packed float4 fields).
template <bool POW2, uint32_t S>
__device__ inline void dda_setup(const GridDev& g, Vec3f s, Vec3f e, Dda* d) {
  int ixe = (int)((e.x - g.xmin) / g.dx);
  int iye = (int)((e.y - g.ymin) / g.dy);
  int ize = (int)((e.z - g.zmin) / g.dz);
  ixe = max(min(ixe, g.nx - 1), 0);
  iye = max(min(iye, g.ny - 1), 0);
  ize = max(min(ize, g.nz - 1), 0);
  int ix = (int)((s.x - g.xmin) / g.dx);
  int iy = (int)((s.y - g.ymin) / g.dy);
  int iz = (int)((s.z - g.zmin) / g.dz);
  ix = max(min(ix, g.nx - 1), 0);
  iy = max(min(iy, g.ny - 1), 0);
  iz = max(min(iz, g.nz - 1), 0);
  d->idxe = cell_index(g, ixe, iye, ize) * S;
  d->idx = cell_index(g, ix, iy, iz) * S;

  float dx = e.x - s.x, dy = e.y - s.y, dz = e.z - s.z;
  float dl = (float)__builtin_sqrt((double)(dx * dx + dy * dy + dz * dz) + kWalkTiny);
  float dxy = (float)__builtin_sqrt((double)(dx * dx + dy * dy));
  float cos_t = dz / dl;
  float sin_t = dxy / dl;
  float cos_p = (float)((double)dx / ((double)dxy + kWalkTiny));
  float sin_p = (float)((double)dy / ((double)dxy + kWalkTiny));
  float xo = cos_p * sin_t;
  if ((double)fabsf(xo) < kWalkTiny) xo = (float)((double)xo >= 0.0 ? kWalkTiny : -kWalkTiny);
  float yo = sin_p * sin_t;
  if (fabs((double)yo) < kWalkTiny) yo = (float)((double)yo >= 0.0 ? kWalkTiny : -kWalkTiny);
  float zo = cos_t;
  if (fabs((double)zo) < kWalkTiny) zo = (float)((double)zo >= 0.0 ? kWalkTiny : -kWalkTiny);
  d->xo = xo;
  d->yo = yo;
  d->zo = zo;
  d->rxo = 1.0 / (double)xo;
  d->ryo = 1.0 / (double)yo;
  d->rzo = 1.0 / (double)zo;
  const bool px = (double)xo > 0.0, py = (double)yo > 0.0, pz = (double)zo > 0.0;
  d->sdx = px ? g.dx : -g.dx;
  d->sdy = py ? g.dy : -g.dy;
  d->sdz = pz ? g.dz : -g.dz;
  if constexpr (POW2) {
    d->fx = (px ? (float)ix + 1.0f : (float)ix) * g.dx;
    d->fy = (py ? (float)iy + 1.0f : (float)iy) * g.dy;
    d->fz = (pz ? (float)iz + 1.0f : (float)iz) * g.dz;
    d->lx = (px ? (float)g.nx + 0.5f : -0.5f) * g.dx;
    d->ly = (py ? (float)g.ny + 0.5f : -0.5f) * g.dy;
    d->lz = (pz ? (float)g.nz + 0.5f : -0.5f) * g.dz;
    d->negx = !px;
    d->negy = !py;
    d->negz = !pz;
  } else {
    d->fx = px ? (float)ix + 1.0f : -(float)ix;
    d->fy = py ? (float)iy + 1.0f : -(float)iy;
    d->fz = pz ? (float)iz + 1.0f : -(float)iz;
    d->lx = px ? (float)(g.nx + 1) : 1.0f;
    d->ly = py ? (float)(g.ny + 1) : 1.0f;
    d->lz = pz ? (float)(g.nz + 1) : 1.0f;
  }
  const uint32_t sz = S, sy = (uint32_t)g.nz * S, sx = (uint32_t)g.ny * (uint32_t)g.nz * S;
  d->stx = px ? sx : 0u - sx;
  d->sty = py ? sy : 0u - sy;
  d->stz = pz ? sz : 0u - sz;
  d->xc = s.x - g.xmin;
  d->yc = s.y - g.ymin;
  d->zc = s.z - g.zmin;
}

// The walk continues while the current cell lies inside the mesh and is not
// the end cell (the index is meaningful only when the cell is inside).
template <bool POW2>
__device__ __forceinline__ bool dda_active(const Dda& d) {
  if constexpr (POW2)
    return ((d.fx < d.lx) != d.negx) & ((d.fy < d.ly) != d.negy) & ((d.fz < d.lz) != d.negz) &
           (d.idx != d.idxe);
  else
    return (d.fx < d.lx) & (d.fy < d.ly) & (d.fz < d.lz) & (d.idx != d.idxe);
}

// One step of calc_optical_depth.c:163-179; returns the chord length.
template <bool POW2>
__device__ __forceinline__ float dda_step(Dda* d) {
  const float ddx = (POW2 ? d->fx : d->sdx * d->fx) - d->xc;
  const float ddy = (POW2 ? d->fy : d->sdy * d->fy) - d->yc;
  const float ddz = (POW2 ? d->fz : d->sdz * d->fz) - d->zc;
  const float rx = div_by(ddx, d->rxo);
  const float ry = div_by(ddy, d->ryo);
  const float rz = div_by(ddz, d->rzo);
  const float rmin = fminf(rx, fminf(ry, rz));
  const bool ax = rmin == rx, ay = rmin == ry, az = rmin == rz;
  d->fx += ax ? (POW2 ? d->sdx : 1.0f) : 0.0f;
  d->fy += ay ? (POW2 ? d->sdy : 1.0f) : 0.0f;
  d->fz += az ? (POW2 ? d->sdz : 1.0f) : 0.0f;
  d->idx += (ax ? d->stx : 0u) + (ay ? d->sty : 0u) + (az ? d->stz : 0u);
  d->xc = d->xc + rmin * d->xo;
  d->yc = d->yc + rmin * d->yo;
  d->zc = d->zc + rmin * d->zo;
  return rmin;
}

// calc_optical_depth.c:194-203: length from the last face to the end point.
__device__ __forceinline__ float dda_depth(const GridDev& g, const Dda& d, Vec3f e) {
  const float x = d.xc + g.xmin, y = d.yc + g.ymin, z = d.zc + g.zmin;
  const float ex = e.x - x, ey = e.y - y, ez = e.z - z;
  return (float)__builtin_sqrt((double)(ex * ex + ey * ey + ez * ez));
}

struct WalkOut {
  float tau[3];
  int ncross;
  uint64_t hash;
};

__device__ __forceinline__ void kahan_add(float& s, float& c, float term) {
  const float y = term - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// Fields of a cell.  A Fields functor states the scale kScale of the
// indices it takes (the walk carries the linear cell index times kScale) and
// whether its loads are range-checked (kChecked: an index outside the mesh
// reads zeros or another cell's data, never memory outside the array, so
// the walk need not redirect the loads of lanes that have left the mesh).
constexpr int64_t kNarrowCells = (int64_t)1 << 28;
constexpr int kBufferWord3 = 0x00020000;   // raw 32-bit buffer, gfx9 layout

// Per-cell FP32 densities (nHI, nHeI, nHeII) from the packed float4 array.
// NARROW (fewer than 2^28 cells): indices are byte offsets (cell * 16, below
// 2^32) and the loads are buffer loads against the array's size, which
// return zeros past its end.
template <bool NARROW>
struct PackedFields {
  const float4* __restrict__ nf;
  uint32_t nbytes;   // NARROW: 16 * cells
  static constexpr uint32_t kScale = NARROW ? 16u : 1u;
  static constexpr bool kChecked = NARROW;
  __device__ __forceinline__ float4 operator()(uint32_t idx) const {
    if constexpr (NARROW) {
      typedef uint32_t U3 __attribute__((ext_vector_type(3)));
      const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<float4*>(nf), (short)0, (int)nbytes, kBufferWord3);
      const U3 v = __builtin_amdgcn_raw_buffer_load_b96(r, idx, 0, 0);
      return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), 0.0f);
    } else {
      return nf[idx];
    }
  }
};

// The whole walk of one segment (calc_optical_depth.c:76-227).  fields(idx)
// gives the per-cell FP32 densities (nHI, nHeI, nHeII) formed as the
// reference forms them.  The densities of the next two cells are fetched
// before the current chord is summed (the DDA runs one cell ahead of the
// sums), so two gathers are in flight.  KAHAN selects
// compensated sums; HASH records the cell sequence (FNV-1a over the linear
// cell indices, the end cell last); POW2 selects the face-position form of
// the DDA for power-of-two cell sizes (same results).
template <bool KAHAN, bool HASH, bool POW2, typename Fields>
__device__ inline void walk_segment_with(const GridDev& g, const Fields& fields, Vec3f s, Vec3f e,
                                         WalkOut* out) {
  constexpr uint32_t S = Fields::kScale;
  // index of a lane's next load: any index when the loads are range-checked,
  // else a cell inside the mesh (cell 0 for a lane whose walk has ended)
  auto at = [&](bool act, uint32_t i) { return (Fields::kChecked || act) ? i : 0u; };
  Dda d;
  dda_setup<POW2, S>(g, s, e, &d);
  float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
  int nc = 0;
  uint64_t h = 0xcbf29ce484222325ull;
  // The loop is unrolled by two so that the two field slots (A, B) alternate
  // without register copies (a copy of a load's result would wait for the
  // load).  Every pass issues both loads, whichever lanes are active, so
  // that the wait before each sum is for the older load only.  Slot A holds
  // cell k (always active inside the loop), slot B cell k+1; the chords are
  // summed in cell order, so the sums are those of the one-slot loop.
  bool actA = dda_active<POW2>(d);
  float4 nA = fields(at(actA, d.idx));
  float rA = 0.0f, rB = 0.0f;
  bool actB = false;
  if (actA) {
    if (HASH) h = (h ^ (uint64_t)(d.idx / S)) * 0x100000001b3ull;
    rA = dda_step<POW2>(&d);
    actB = dda_active<POW2>(d);
  }
  float4 nB = fields(at(actB, d.idx));
  auto add = [&](float r, const float4& n) {
    if (KAHAN) {
      kahan_add(t0, c0, r * n.x);
      kahan_add(t1, c1, r * n.y);
      kahan_add(t2, c2, r * n.z);
    } else {
      t0 = t0 + r * n.x;
      t1 = t1 + r * n.y;
      t2 = t2 + r * n.z;
    }
  };
  if (actA) do {
    nc++;
    add(rA, nA);
    bool a2 = false;
    if (actB) {
      if (HASH) h = (h ^ (uint64_t)(d.idx / S)) * 0x100000001b3ull;
      rB = dda_step<POW2>(&d);
      a2 = dda_active<POW2>(d);
    }
    nA = fields(at(a2, d.idx));
    bool a3 = false;
    if (actB) {
      nc++;
      add(rB, nB);
      if (a2) {
        if (HASH) h = (h ^ (uint64_t)(d.idx / S)) * 0x100000001b3ull;
        rA = dda_step<POW2>(&d);
        a3 = dda_active<POW2>(d);
      }
    }
    nB = fields(at(a3, d.idx));
    actA = a2;
    actB = a3;
  } while (actA);
  const float depth = dda_depth(g, d, e);
  const uint32_t ie = d.idxe;
  float4 n = fields(ie);
  if (HASH) h = (h ^ (uint64_t)(ie / S)) * 0x100000001b3ull;
  nc++;
  if (KAHAN) {
    kahan_add(t0, c0, depth * n.x);
    kahan_add(t1, c1, depth * n.y);
    kahan_add(t2, c2, depth * n.z);
  } else {
    t0 = t0 + depth * n.x;
    t1 = t1 + depth * n.y;
    t2 = t2 + depth * n.z;
  }
  out->tau[0] = t0;
  out->tau[1] = t1;
  out->tau[2] = t2;
  out->ncross = nc;
  out->hash = h;
}

// The walk over the packed per-cell densities nf (float4 per cell).
template <bool KAHAN, bool HASH, bool NARROW, bool POW2>
__device__ __forceinline__ void walk_segment(const GridDev& g, const float4* __restrict__ nf,
                                             Vec3f s, Vec3f e, WalkOut* out) {
  const uint32_t nbytes = NARROW ? (uint32_t)((int64_t)g.nx * g.nyz * 16) : 0u;
  walk_segment_with<KAHAN, HASH, POW2>(g, PackedFields<NARROW>{nf, nbytes}, s, e, out);
}

// delta_length, calc_photoion_rate.c:52-90.
__device__ inline float delta_length(const GridDev& g, float dx, float dy, float dz) {
  float dl = (float)__builtin_sqrt((double)(dx * dx + dy * dy + dz * dz) + kRateTiny);
  float dxy = (float)__builtin_sqrt((double)(dx * dx + dy * dy) + kRateTiny);
  float cos_t = dz / dl;
  float sin_t = dxy / dl;
  float cos_p = dx / dxy;
  float sin_p = dy / dxy;
  float xo = cos_p * sin_t;
  if ((double)fabsf(xo) < kRateTiny) xo = (float)((double)xo >= 0.0 ? kRateTiny : -kRateTiny);
  float yo = sin_p * sin_t;
  if ((double)fabsf(yo) < kRateTiny) yo = (float)((double)yo >= 0.0 ? kRateTiny : -kRateTiny);
  float zo = cos_t;
  if ((double)fabsf(zo) < kRateTiny) zo = (float)((double)zo >= 0.0 ? kRateTiny : -kRateTiny);
  const float rx = fabsf(g.dx / xo);
  const float ry = fabsf(g.dy / yo);
  const float rz = fabsf(g.dz / zo);
  return fminf(rx, fminf(ry, rz));
}

// Per-ray quantities of calc_photoion_rate.c:159-246 that do not depend on
// frequency.
struct RayRate {
  double t[3];    // column densities at the cell entry [cm^-2]
  double d[3];    // in-cell optical-depth weights n_s*delta*lunit
  double r[3];    // 1/(n_s V_sh + TINY)
};

__device__ inline void ray_rate_setup(const GridDev& g, const RateTables& rt, const float tau[3],
                                      Vec3f src, Vec3f tgt, float dens, float fHI, float fHeI,
                                      float fHeII, RayRate* rr) {
  const double col = g.lunit * g.denstonh;
  rr->t[0] = (double)tau[0] * col;
  rr->t[1] = (double)tau[1] * col;
  rr->t[2] = (double)tau[2] * col;
  const float dx = tgt.x - src.x, dy = tgt.y - src.y, dz = tgt.z - src.z;
  const float delta = delta_length(g, dx, dy, dz);
  const float dist = (float)__builtin_sqrt((double)(dx * dx + dy * dy + dz * dz));
  const double nH = (double)dens * g.denstonh;
  const double nHI = nH * (double)fHI;
  rr->d[0] = nHI * (double)delta * g.lunit;
  rr->t[0] -= 0.5 * rr->d[0];
  const double nHe = nH * g.helium_fact;
  const double nHeI = nHe * (double)fHeI;
  const double nHeII = nHe * (double)fHeII;
  rr->d[1] = nHeI * (double)delta * g.lunit;
  rr->d[2] = nHeII * (double)delta * g.lunit;
  rr->t[1] -= 0.5 * rr->d[1];
  rr->t[2] -= 0.5 * rr->d[2];
  const double a = (double)dist + 0.5 * (double)delta;
  const double b = (double)dist - 0.5 * (double)delta;
  const double lu = g.lunit;
  const double vsh = 4.0 * kPi / 3.0 * (a * a * a - b * b * b) * (lu * lu * lu);
  rr->r[0] = 1.0 / (nHI * vsh + kRateTiny);
  rr->r[1] = 1.0 / (nHeI * vsh + kRateTiny);
  rr->r[2] = 1.0 / (nHeII * vsh + kRateTiny);
}

// The reference's frequency loop in FP64, calc_photoion_rate.c:285-387,
// bin by bin (used when FP32 would overflow, e.g. a target cell that holds
// the ray's source, where the reference itself produces inf/NaN).
template <typename SpecFn>
__device__ inline void rate_loop_f64(const RateTables& rt, const RayRate& rr, SpecFn spec,
                                     double out[6]) {
  double g0 = 0, h0 = 0, g1 = 0, g2 = 0, h1 = 0, h2 = 0;
  for (int k = 0; k < rt.ngrid; k++) {
    const float nu = rt.nu[k];
    const double s0 = rt.cs[3 * k + 0], s1 = rt.cs[3 * k + 1], s2 = rt.cs[3 * k + 2];
    double tau = rr.t[0] * s0;
    double w0 = rr.d[0] * s0;
    double dtot = w0;
    tau += rr.t[1] * s1;
    tau += rr.t[2] * s2;
    double w1 = rr.d[1] * s1;
    double w2 = rr.d[2] * s2;
    dtot += w1 + w2;
    const double absorbed = spec(k) * exp(-tau) * (1.0e0 - exp(-dtot));
    w0 /= (dtot + kRateTiny);
    g0 += w0 * absorbed * rr.r[0];
    if ((double)nu > 1.0) h0 += w0 * absorbed * ((double)nu - 1.0) * rr.r[0];
    w1 /= (dtot + kRateTiny);
    w2 /= (dtot + kRateTiny);
    g1 += w1 * absorbed * rr.r[1];
    g2 += w2 * absorbed * rr.r[2];
    if ((double)nu > 1.8125) h1 += w1 * absorbed * ((double)nu - 1.8125) * rr.r[1];
    if ((double)nu > 4.0) h2 += w2 * absorbed * ((double)nu - 4.0) * rr.r[2];
  }
  out[0] = g0;
  out[1] = h0 * kHnuL;
  out[2] = g1;
  out[3] = g2;
  out[4] = h1 * kHnuL;
  out[5] = h2 * kHnuL;
}

}  // namespace argot
