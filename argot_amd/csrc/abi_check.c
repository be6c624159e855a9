/* Compile-time check that the C-ABI structs have the reference's layout
 * (gcc 11 x86-64, __HELIUM__ on, __HYDROGEN_MOL__ off, NGRID_NU = 256). */
#include <stddef.h>
#include "argot_hip.h"

_Static_assert(sizeof(struct prim_chem) == 52, "prim_chem");
_Static_assert(offsetof(struct prim_chem, fHI) == 8, "prim_chem.fHI");
_Static_assert(offsetof(struct prim_chem, fHeI) == 16, "prim_chem.fHeI");
_Static_assert(offsetof(struct prim_chem, GammaHI) == 28, "prim_chem.GammaHI");
_Static_assert(offsetof(struct prim_chem, HeatHeII) == 48, "prim_chem.HeatHeII");
_Static_assert(sizeof(struct fluid_mesh) == 156, "fluid_mesh");
_Static_assert(offsetof(struct fluid_mesh, chem) == 44, "fluid_mesh.chem");
_Static_assert(offsetof(struct fluid_mesh, prev_chem) == 96, "fluid_mesh.prev_chem");
_Static_assert(offsetof(struct fluid_mesh, prev_uene) == 148, "fluid_mesh.prev_uene");
_Static_assert(sizeof(struct freq_param) == 2072, "freq_param");
_Static_assert(sizeof(struct radiation_src) == 2072, "radiation_src");
_Static_assert(offsetof(struct radiation_src, photon_rate) == 24, "radiation_src.photon_rate");
_Static_assert(sizeof(struct cross_section) == 12, "cross_section");
_Static_assert(sizeof(struct run_param) == 5480, "run_param");
_Static_assert(offsetof(struct run_param, lunit) == 80, "run_param.lunit");
_Static_assert(offsetof(struct run_param, denstonh) == 112, "run_param.denstonh");
_Static_assert(offsetof(struct run_param, freq) == 184, "run_param.freq");
_Static_assert(offsetof(struct run_param, csect) == 2256, "run_param.csect");
_Static_assert(offsetof(struct run_param, nray) == 5368, "run_param.nray");
_Static_assert(offsetof(struct run_param, nseg) == 5376, "run_param.nseg");
_Static_assert(offsetof(struct run_param, xmin_local) == 5408, "run_param.xmin_local");
_Static_assert(offsetof(struct run_param, delta_x) == 5456, "run_param.delta_x");
_Static_assert(sizeof(struct photoion_rate) == 24, "photoion_rate");
_Static_assert(sizeof(struct ray_segment) == 56, "ray_segment");
_Static_assert(offsetof(struct ray_segment, xpos_start) == 16, "ray_segment.start");
_Static_assert(offsetof(struct ray_segment, optical_depth_HI) == 40, "ray_segment.tau");
_Static_assert(sizeof(struct light_ray) == 2384, "light_ray");
_Static_assert(sizeof(struct light_ray_IO) == 2096, "light_ray_IO");
_Static_assert(offsetof(struct light_ray_IO, ix_target) == 2072, "light_ray_IO.ix_target");
_Static_assert(offsetof(struct light_ray_IO, optical_depth_HI) == 2084, "light_ray_IO.tau");
_Static_assert(sizeof(struct cuda_param) == 88, "cuda_param");
_Static_assert(sizeof(struct cuda_mem_space) == 88, "cuda_mem_space");

int argot_hip_abi_checked(void) { return 1; }
