// Rate surrogate: device view used by the fused step kernel, and the host
// hooks the runtime calls.  The network, its training and validation live
// in surrogate.hip.
#pragma once
#include <vector>

#include "mlp.h"

namespace argot {

// Everything the fused kernel needs to serve a ray from the network.
struct SurrDev {
  MlpDev net;
  const unsigned char* trust;  // [n^3] 1 where the validated error is within tolerance
  int trust_n;                 // trust cells per axis
  float lo, hi;                // input hull, log10 threshold optical depth
  float inv_cell;              // trust_n / (hi - lo)
  float sig_th[3];             // threshold cross-sections: tau_th,s = N_s * sig_th[s]
  float dtau_max;              // largest in-cell optical depth for the 2-point rule
  double tau_lo, tau_hi;       // 10^lo, 10^hi
  const double* node_q;        // [nnode] photon rate of a node whose spectrum has the
                               // trained shape, else -1
};

// Host hooks (surrogate.hip).
bool surrogate_ready();
// Trained and at least one trust cell within tolerance.
bool surrogate_any_trusted();
// Status of the calling thread's last entry point (runtime.hip).
void argot_hip_internal_set_error(int code);
void argot_hip_internal_clear_error();
// n_He/n_H of argot_hip_set_helium_fact (runtime.hip).
double argot_hip_internal_helium_fact();
// Device view (node_q left null; the runtime fills it in).
SurrDev surrogate_dev();
// For every tree node (spectra spec64[nnode*ngrid]): its total photon rate
// if the spectrum has the trained shape to 1e-9 per bin, else -1.
void surrogate_node_q(const double* spec64, int nnode, int ngrid, std::vector<double>* q);

}  // namespace argot
