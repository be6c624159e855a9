// Host-side source tree of libargot_hip.so (see tree.cpp).
#pragma once
#include <stdint.h>

#include <vector>

#include "argot_internal.h"

namespace argot {

struct TreeBox {
  float xmin, ymin, zmin, xmax, ymax, zmax;
};

struct HostTree {
  int ngrid = 0, nsrc = 0;
  float theta_sq = 0.0f;
  std::vector<NodeLink> link;      // pre-order
  std::vector<float> cm;           // [4*nnode] cm xyz, SQR(length)
  std::vector<double> spec64;      // [nnode*ngrid]
  std::vector<float> spec;         // [nnode*ngrid] normalised by spec_scale
  std::vector<double> spec_scale;  // [nnode]
  std::vector<float> src_pos;      // [4*nsrc]
  int nnode() const { return (int)link.size(); }
};

struct HostRay {
  int kind;   // 0 node, 1 source
  int id;     // node (pre-order) or source index
  float pos[3];
};

// Positions pos[3*i] and spectra spec[i*spec_stride + k] (k < ngrid) of nsrc
// sources; returns 0 on success.
int build_tree(const float* pos, const double* spec, size_t spec_stride, int nsrc, int ngrid,
               const TreeBox& box, float theta_crit, HostTree* t);

// Rays for one target point (FP32 opening test as the kernel does it).
uint64_t tree_rays_to(const HostTree& t, const float* target, std::vector<HostRay>* out);

}  // namespace argot
