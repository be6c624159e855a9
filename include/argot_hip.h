/*
 * argot_hip.h — C-ABI of libargot_hip.so, the MI355X (gfx950) implementation
 * of ARGOT's point-source radiation step.
 *
 * Drop-in boundary: the structs below have the byte layout of the
 * reference's own (src/radiation.h, src/fluid.h, src/run_param.h,
 * src/source.h, src/cross_section.h, src/cuda_mem_space.h) for the build
 * the benchmarks use: __HELIUM__ on, __HYDROGEN_MOL__ off, NGRID_NU = 256,
 * NSEG_PER_RAY = 5, NMAX_CUDA_DEV = 4, x86-64.  Sizes and offsets are
 * checked with _Static_assert in argot_amd/csrc/abi_check.c.  The entry
 * points keep the reference's names, argument lists and void returns
 * (src/prototype.h:86-130); failures are reported through
 * argot_hip_last_error() instead of assert/exit.
 *
 * Mesh sizes are read at run time from run_param.nmesh_{x,y,z}_local (the
 * reference fixes them at compile time in run_param.h:22-74).
 */
#ifndef ARGOT_HIP_H
#define ARGOT_HIP_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARGOT_NGRID_NU 256
#define ARGOT_NSEG_PER_RAY 5
#define ARGOT_NMAX_DEV 4

/* A stream handle: the reference's C side sees cudaStream_t as int64_t
 * (src/cuda_mem_space.h:10-12); here it is a hipStream_t (8 bytes). */
typedef void *argot_stream_t;

/* src/prim_chem/chemistry.h (52 B with __HELIUM__) */
struct prim_chem {
    float wmol, felec;
    float fHI, fHII;
    float fHeI, fHeII, fHeIII;
    float GammaHI, HeatHI;
    float GammaHeI, GammaHeII;
    float HeatHeI, HeatHeII;
};

/* src/fluid.h:123-138 (156 B) */
struct fluid_mesh {
    float dens, eneg, momx, momy, momz;
    float uene, duene, durad, etrp, pot, dens_prev;
    struct prim_chem chem;
    struct prim_chem prev_chem;
    float prev_uene;
    short under_shock, high_mach;
};

/* src/source.h:6-16 */
struct freq_param {
    double log_nu_min, log_nu_max, dlog_nu;
    double nu[ARGOT_NGRID_NU];
};

struct radiation_src {
    int type;                 /* 0: blackbody, 1: power law */
    float param;              /* temperature [K] or index */
    float xpos, ypos, zpos;
    double photon_rate[ARGOT_NGRID_NU];
};

/* src/cross_section.h:7-20 */
struct cross_section {
    float csect_HI, csect_HeI, csect_HeII;
};

/* src/cosmology.h:8-11 */
struct cosmology {
    float omega_m, omega_v, omega_b, omega_nu, hubble;
    float tend;
};

/* src/run_param.h:171-235 (5480 B) */
struct run_param {
    char *model_name;
    FILE *proc_file, *diag_file, *err_file;
    int helium_flag, hydrogen_mol_flag;
    int cosmology_flag;
    int ngrid_nu;
    int mpi_rank, mpi_nproc;
    int rank_x, rank_y, rank_z;
    int nnode_x, nnode_y, nnode_z;
    double lunit, munit, tunit, eunit;
    double denstonh, uenetok, masstonh;
    float anow, tnow, znow, hnow;
    struct cosmology cosm;
    float tend;
    float dtime;
    struct freq_param freq;
    struct cross_section csect[ARGOT_NGRID_NU];
    int nspecies;
    int nchannel;
    int step;
    int output_indx;
    int noutput;
    float *output_timing;
    uint64_t nsrc;
    uint64_t nray;
    uint64_t nseg;
    float xmin, ymin, zmin;
    float xmax, ymax, zmax;
    float xmin_local, ymin_local, zmin_local;
    float xmax_local, ymax_local, zmax_local;
    int nmesh_x_total, nmesh_y_total, nmesh_z_total;
    int nmesh_x_local, nmesh_y_local, nmesh_z_local;
    float delta_x, delta_y, delta_z;
    float theta_crit;
    int nmesh_per_loop;
};

/* src/radiation.h:10-83 */
struct photoion_rate {
    float GammaHI, HeatHI;
    float GammaHeI, GammaHeII;
    float HeatHeI, HeatHeII;
};

struct ray_segment {
    uint64_t ray_indx;
    int local_rank;
    int target_rank;
    float xpos_start, ypos_start, zpos_start;
    float xpos_end, ypos_end, zpos_end;
    float optical_depth_HI;
    float optical_depth_HeI, optical_depth_HeII;
};

struct light_ray {
    struct radiation_src src;
    int num_segment;
    struct ray_segment segment[ARGOT_NSEG_PER_RAY];
    int ix_target, iy_target, iz_target;
    float optical_depth_HI;
    float optical_depth_HeI, optical_depth_HeII;
};

struct light_ray_IO {
    struct radiation_src src;
    int ix_target, iy_target, iz_target;
    float optical_depth_HI;
    float optical_depth_HeI, optical_depth_HeII;
};

/* src/cuda_mem_space.h:14-39 */
struct cuda_param {
    int num_cuda_dev;
    int nmesh_per_dev;
    int cuda_nblock;
    int max_thread_dimx;
    int max_block_dimx;
    argot_stream_t strm[ARGOT_NMAX_DEV];
    argot_stream_t diffuse_strm[ARGOT_NMAX_DEV];
};

struct cuda_mem_space {
    struct fluid_mesh *mesh_dev;
    struct run_param *this_run_dev;
    struct freq_param *freq_dev;
    struct ray_segment *segment_dev;
    struct photoion_rate *rate_dev;
    struct prim_chem *chem_updated;
    struct light_ray_IO *ray_IO_dev;
    float *optical_depth_dev;
    float *dtime_dev;
    float *diff_chem_dev;
    float *diff_uene_dev;
};

/* Only passed through (src/mpi_param.h); one rank needs nothing from it. */
struct mpi_param;

/* ------------------------------------------------------------------ */
/* Table set-up (host only)                                            */
/* ------------------------------------------------------------------ */

/* replaces setup_freq_param, src/source.c:38-53 */
void setup_freq_param(struct freq_param *freq);
/* replaces setup_photon_rate, src/source.c:104-138 */
void setup_photon_rate(struct freq_param *freq, struct radiation_src *src,
                       double total_photon_rate);
/* replaces setup_cross_section, src/cross_section.c:5-22 */
void setup_cross_section(struct cross_section *csect, struct freq_param *freq);

/* ------------------------------------------------------------------ */
/* GPU entry points (src/prototype.h:86-130)                           */
/* ------------------------------------------------------------------ */

/* replaces init_gpu, src/init_gpu.cu:14-145: one device (num_cuda_dev=1),
 * device-resident mesh (AoS, as the reference) and run_param, stream. */
void init_gpu(struct fluid_mesh *mesh, struct cuda_mem_space *cuda_mem,
              struct cuda_param *this_cuda, struct run_param *this_run);
/* replaces send_mesh_data, src/send_mesh_data.cu:8-29 */
void send_mesh_data(struct fluid_mesh *mesh, struct cuda_mem_space *cuda_mem,
                    struct cuda_param *this_cuda, struct run_param *this_run);
/* replaces recv_mesh_data, src/recv_mesh_data.cu:8-34 */
void recv_mesh_data(struct fluid_mesh *mesh, struct cuda_mem_space *cuda_mem,
                    struct cuda_param *this_cuda, struct run_param *this_run);
/* replaces send_run_param_data, src/send_run_param_data.cu:7-21 */
void send_run_param_data(struct run_param *this_run,
                         struct cuda_mem_space *cuda_mem,
                         struct cuda_param *this_cuda);
/* replaces allocate_pinned_segment etc., src/allocate_pinned_segment.cu:14-37 */
void allocate_pinned_segment(struct ray_segment **seg, uint64_t nseg);
void deallocate_pinned_segment(struct ray_segment *seg);
void allocate_pinned_light_ray_IO(struct light_ray_IO **ray_IO, uint64_t nray);
void deallocate_pinned_light_ray_IO(struct light_ray_IO *ray_IO);

/* replaces calc_optical_depth_cuda, src/calc_optical_depth_cuda.cu:29-144:
 * this_run->nseg host segments in, their optical depths overwritten. */
void calc_optical_depth_cuda(struct ray_segment *seg,
                             struct cuda_mem_space *cuda_mem,
                             struct cuda_param *this_cuda,
                             struct run_param *this_run);
/* replaces calc_photoion_rate_cuda, src/calc_photoion_rate_cuda.cu:92-216:
 * this_run->nray rays; ray_IO is the caller's pinned staging buffer; the
 * rates are added into mesh_dev[].prev_chem (no atomics: per-target
 * segmented sums in ray order). */
void calc_photoion_rate_cuda(struct cuda_mem_space *cuda_mem,
                             struct light_ray *ray,
                             struct light_ray_IO *ray_IO,
                             struct cuda_param *this_cuda,
                             struct run_param *this_run);
/* replaces zero_out_photoion_rate_cuda, src/calc_photoion_rate_cuda.cu:45-62 */
void zero_out_photoion_rate_cuda(struct cuda_mem_space *cuda_mem,
                                 struct cuda_param *this_cuda,
                                 struct run_param *this_run);
/* replaces merge_photoion_data, src/merge_photoion_data.cu:16-83:
 * mesh_dev[].prev_chem += host mesh[].prev_chem rates. */
void merge_photoion_data(struct fluid_mesh *mesh,
                         struct cuda_mem_space *cuda_mem,
                         struct cuda_param *this_cuda,
                         struct run_param *this_run);

/* replaces step_radiation_tree (GPU variant), src/step_radiation_tree.c:554-733
 * and src/prototype.h:97: tree from src[0..this_run->nsrc), then for every
 * target cell the rays, segments, walk, rates and per-cell sums, fused on
 * the GPU; rates land in mesh_dev[].prev_chem (zeroed first).  Smoothing
 * and the chemistry iteration are outside this library: one pass. */
void step_radiation_tree(struct fluid_mesh *mesh, struct radiation_src *src,
                         struct run_param *this_run, struct mpi_param *this_mpi,
                         struct cuda_mem_space *cuda_mem,
                         struct cuda_param *this_cuda, float dtime);

/* ------------------------------------------------------------------ */
/* Library controls (no reference counterpart)                         */
/* ------------------------------------------------------------------ */

enum argot_hip_status {
    ARGOT_OK = 0,
    ARGOT_ERR_HIP = 1,          /* a HIP runtime call failed */
    ARGOT_ERR_ARG = 2,          /* bad argument or size */
    ARGOT_ERR_STATE = 3,        /* init_gpu not called / mesh not sent */
    ARGOT_ERR_NOMEM = 4,        /* device allocation failed */
    ARGOT_ERR_TOO_MANY_SEGMENTS = 5  /* ray crosses too many domains */
};

/* Status of the last entry point called on this thread (0 = OK). */
int argot_hip_last_error(void);
const char *argot_hip_error_string(int code);

/* n_He/n_H used for the helium number densities (reference: the
 * compile-time HELIUM_FACT, chemistry.h:25, default 0.0811258278). */
void argot_hip_set_helium_fact(double helium_fact);

/* Optical-depth summation: 0 = the reference's plain FP32 sums (default),
 * 1 = Kahan-compensated FP32 sums. */
void argot_hip_set_tau_mode(int mode);

/* Rate integral: 0 = exact frequency integral (default), 1 = MLP
 * surrogate with exact fallback outside its training range. */
void argot_hip_set_rate_mode(int mode);

/* Counters of the last step_radiation_tree call. */
struct argot_step_stats {
    uint64_t ntarget;        /* target cells */
    uint64_t nray;           /* rays (== segments, one rank) */
    uint64_t ncross;         /* cell crossings incl. the end-cell term */
    uint64_t nfallback;      /* surrogate rays that used the exact integral */
    double ms_build;         /* host tree build + upload */
    double ms_kernel;        /* fused kernel(s), HIP events */
    double ms_total;         /* whole call, wall clock */
};
void argot_hip_step_stats(struct argot_step_stats *out);

/* Opening angle for step_radiation_tree (default 0.7 = ARGOT_THETA_CRIT). */
void argot_hip_set_theta(float theta);

/* Persistent grid of the fused kernel, in wavefronts per compute unit. */
void argot_hip_set_blocks_per_cu(int n);

/* Measured device copy bandwidth: best of reps copies of nbytes with a
 * plain streaming kernel, in GB/s of bytes read plus written. */
double argot_hip_copy_probe(int64_t nbytes, int reps);

#ifdef __cplusplus
}
#endif
#endif /* ARGOT_HIP_H */
