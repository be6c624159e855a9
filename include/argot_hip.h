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
 * device-resident mesh (AoS, as the reference) and run_param, stream.
 * ARGOT_ERR_ARG for a local mesh with a non-positive size, nx*ny >= 2^24 or
 * more than 2^31-1 cells.
 * Each call creates a context of its own (device mesh, stream, buffers,
 * step graph) that every later call finds through cuda_mem->mesh_dev: two
 * meshes, each with its own cuda_mem_space, coexist in one process, and
 * calls on different cuda_mem_spaces may run concurrently from different
 * threads (calls on one cuda_mem_space must not overlap).  Calling init_gpu
 * again on a cuda_mem_space that names a context frees that context first.
 * The library controls below (argot_hip_set_*) are process-wide. */
void init_gpu(struct fluid_mesh *mesh, struct cuda_mem_space *cuda_mem,
              struct cuda_param *this_cuda, struct run_param *this_run);
/* No counterpart in the reference (its process keeps its device memory
 * until it ends): frees the context init_gpu made for cuda_mem and zeroes
 * cuda_mem.  Returns ARGOT_OK, or ARGOT_ERR_STATE if cuda_mem names no
 * context; argot_hip_last_error() is left as it was. */
int argot_hip_release_gpu(struct cuda_mem_space *cuda_mem);
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
/* No counterpart in the reference (whose calc_optical_depth_cuda allocates
 * and frees its device segment buffer on every call,
 * src/calc_optical_depth_cuda.cu:69,131).  For callers that cannot
 * allocate device memory: serving loops, code running between HIP graph
 * captures, harnesses that forbid hipMalloc.  calc_optical_depth_cuda's
 * semantics and optical-depth bits, with no hipMalloc, hipFree or other
 * allocating call: the segments are staged in a fixed device array of
 * ARGOT_NOALLOC_MAX_SEGMENTS records (56 MiB, one per GPU, held until the
 * process ends) and the densities are formed from mesh_dev's records.
 * this_run->nseg above that limit gives ARGOT_ERR_ARG.  Every call on one
 * GPU shares that array; the library serialises such calls (each holds a
 * lock until its segments are copied back), and the array keeps the last
 * call's segments. */
#define ARGOT_NOALLOC_MAX_SEGMENTS (1 << 20)
void argot_hip_calc_optical_depth_noalloc(struct ray_segment *seg,
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
/* replaces smooth_photoion_rate (GPU variant), src/prototype.h:89 and
 * src/smooth_photoion_rate.c:27-1588, on one rank: the 27-point smoothing
 * (central weight 30, ghost cells = nearest interior cell) of the rates in
 * mesh_dev[].prev_chem, on the device; the smoothed rates are also written
 * into the host mesh's prev_chem when mesh != NULL, as the reference does. */
void smooth_photoion_rate(struct fluid_mesh *mesh, struct run_param *this_run,
                          struct cuda_mem_space *cuda_mem, struct cuda_param *this_cuda,
                          struct mpi_param *this_mpi);

/* Chemistry on the device (helium on, H2 off, heating/cooling on, case-B
 * recombination), one thread per cell.  Each cell starts from
 * mesh_dev[].prev_chem (fractions and the photoionization/heating rates)
 * and prev_uene and integrates the network and the energy over dtime (code
 * units); the results go to mesh_dev[].chem and .uene, .duene := 0.
 * Units come from this_run (lunit, munit, tunit, uenetok, denstonh, anow,
 * znow); the helium fraction from argot_hip_set_helium_fact. */
/* replaces step_chemistry_gpu, src/chemistry.cu:125-222: *diff_chem and
 * *diff_uene receive max(FLT_MIN, the largest relative change over cells)
 * of (GammaHI, felec) against the previous mesh_dev[].chem, and of uene. */
void step_chemistry_gpu(struct cuda_mem_space *cuda_mem, struct cuda_param *this_cuda,
                        struct run_param *this_run, float *diff_chem, float *diff_uene,
                        float dtime);
/* replaces step_chemistry (GPU variant), src/prototype.h:92 and
 * src/step_chemistry.c:53-88: step_chemistry_gpu on one rank (the
 * reduction over ranks is the identity). */
void step_chemistry(struct cuda_mem_space *cuda_mem, struct cuda_param *this_cuda,
                    struct run_param *this_run, float dtime, float *max_diff_chem,
                    float *max_diff_uene);
/* replaces calc_dtime_gpu, src/chemistry.cu:56-120: the smallest safe
 * chemistry step over cells, in seconds. */
double calc_dtime_gpu(struct cuda_mem_space *cuda_mem, struct cuda_param *this_cuda,
                      struct run_param *this_run);
/* replaces calc_timestep_chem (GPU variant), src/prototype.h:90 and
 * src/calc_timestep.c:21-32: calc_dtime_gpu / tunit on one rank. */
double calc_timestep_chem(struct cuda_mem_space *cuda_mem, struct cuda_param *this_cuda,
                          struct run_param *this_run);
/* replaces update_chemistry_gpu, src/prototype.h:118 and
 * src/chemistry.cu:32-54: prev_chem := chem, prev_uene := uene,
 * eneg := uene dens + |mom|^2 / (2 dens). */
void update_chemistry_gpu(struct cuda_mem_space *cuda_mem, struct cuda_param *this_cuda,
                          struct run_param *this_run);

/* On-disk formats (host only; no GPU needed).  Mesh file
 * "<prefix>_XXX_YYY_ZZZ" (rank coordinates): struct run_param as it lies in
 * memory, then one struct fluid_mesh_io (dens eneg momx momy momz uene pot
 * chem, 80 B) per local cell in mesh order.  Source file "<prefix>_src.dat":
 * uint64 nsrc, struct freq_param, nsrc struct radiation_src.
 * The reader checks the file's header against the caller's run_param:
 * nmesh_*_local must equal the caller's (the mesh it allocated); the global
 * mesh and rank counts must equal the caller's when those are set; nspecies
 * and nchannel must be this build's (5, 3).  A mismatch gives ARGOT_ERR_ARG
 * (the reference exits), an unreadable or short file ARGOT_ERR_IO; on
 * error nothing is written to mesh, src or this_run. */
/* replaces input_mesh_single, src/input_data.c:117-150: this_run := the
 * file's header (pointer fields included, as the reference reads them);
 * mesh[].{dens,eneg,mom*,uene,pot,chem} from the records, prev_chem := chem,
 * prev_uene := uene; other fields untouched. */
void input_mesh_single(struct fluid_mesh *mesh, struct run_param *this_run, char *filename);
/* replaces input_mesh, src/input_data.c:152-162 (file "<prefix>_XXX_YYY_ZZZ") */
void input_mesh(struct fluid_mesh *mesh, struct run_param *this_run, char *prefix);
/* replaces input_src_file, src/input_data.c:192-217: this_run->nsrc and
 * ->freq from the file, src[0..nsrc) from its records; nsrc above
 * argot_hip_set_max_sources gives ARGOT_ERR_ARG (reference: NSOURCE_MAX). */
void input_src_file(struct radiation_src *src, struct run_param *this_run, char *src_file);
/* replaces input_src, src/input_data.c:164-190 (file "<prefix>_src.dat") */
void input_src(struct radiation_src *src, struct run_param *this_run, char *prefix);
/* replaces input_data, src/input_data.c:220-230 (no cosmology) */
void input_data(struct fluid_mesh *mesh, struct radiation_src *src, struct run_param *this_run,
                char *prefix);
/* replaces output_mesh_single / output_mesh, src/output_data.c:14-57 */
void output_mesh_single(struct fluid_mesh *mesh, struct run_param *this_run, char *filename);
void output_mesh(struct fluid_mesh *mesh, struct run_param *this_run, char *prefix);
/* replaces output_src / output_data, src/output_data.c:60-86 */
void output_src(struct radiation_src *src, struct run_param *this_run, char *prefix);
void output_data(struct fluid_mesh *mesh, struct radiation_src *src, struct run_param *this_run,
                 char *prefix);
/* Capacity of the caller's source array for input_src (default 1024,
 * NSOURCE_MAX of src/run_param.h:74). */
void argot_hip_set_max_sources(uint64_t n);

/* Counters of the last step_chemistry(_gpu) call on this thread. */
struct argot_chem_stats {
    uint64_t ncell;          /* cells */
    uint64_t nsub;           /* network sub-steps, summed over cells */
    uint64_t niter;          /* energy bisections, summed over cells */
    double ms_kernel;        /* chemistry kernel, HIP events */
    uint64_t nstopped;       /* cells stopped before dtime: their state turned
                                NaN (e.g. an infinite rate in a cell holding a
                                source), or they reached 1e6 sub-steps; the
                                reference would keep taking 0.1-s steps */
};
void argot_hip_chem_stats(struct argot_chem_stats *out);

/* Iterations of step_radiation_tree.  0 (the default): one point-source
 * pass, rates left in mesh_dev[].prev_chem.  n > 0: the reference's loop
 * (src/step_radiation_tree.c:666-723) of at most n passes of transfer,
 * smooth_photoion_rate and step_chemistry(dtime) while the largest relative
 * change of chemistry or energy exceeds 1e-3, then update_chemistry_gpu;
 * the reference runs n = 5. */
void argot_hip_set_iterations(int n);
/* n >= 1: the chemistry kernel is persistent and takes cells from a queue
 * as its lanes finish, so cells with many sub-steps do not hold up a
 * wavefront; free lanes take new cells once n of them are free (or none is
 * busy), n at most 64; 0: one cell per thread.  Results are identical for
 * every n. */
void argot_hip_set_chem_queue(int n);
/* Host threads calc_photoion_rate_cuda uses to read the caller's rays and
 * number their distinct spectra: n >= 1 caps them (default 16; at most one
 * per 4096 rays and per hardware thread); 0 reads and numbers serially.
 * Results are identical. */
void argot_hip_set_host_threads(int n);
/* 1: the walks address the per-cell fields with 64-bit offsets and count
 * cell faces with integer-valued counters on any mesh; 0 (default): 32-bit
 * byte offsets when the mesh has fewer than 2^28 cells (64-bit otherwise),
 * and running face positions instead of counters when every cell size is a
 * power of two.  Results are identical (a test control for the large-mesh
 * and general-spacing paths). */
void argot_hip_set_wide_index(int on);
/* replaces recv_photoion_data, src/recv_photoion_data.cu:14-43:
 * gamma[c] := mesh_dev[c].prev_chem rates (through cuda_mem->rate_dev). */
void recv_photoion_data(struct photoion_rate *gamma, struct cuda_mem_space *cuda_mem,
                        struct cuda_param *this_cuda, struct run_param *this_run);
/* replaces send_photoion_data, src/send_photoion_data.cu:12-41:
 * mesh_dev[c].prev_chem rates := gamma[c]. */
void send_photoion_data(struct photoion_rate *gamma, struct cuda_mem_space *cuda_mem,
                        struct cuda_param *this_cuda, struct run_param *this_run);
/* replaces merge_photoion_data, src/merge_photoion_data.cu:16-83:
 * mesh_dev[].prev_chem += host mesh[].prev_chem rates. */
void merge_photoion_data(struct fluid_mesh *mesh,
                         struct cuda_mem_space *cuda_mem,
                         struct cuda_param *this_cuda,
                         struct run_param *this_run);

/* replaces step_radiation_tree (GPU variant), src/step_radiation_tree.c:554-733
 * and src/prototype.h:97: tree from src[0..this_run->nsrc), then for every
 * target cell the rays, segments, walk, rates and per-cell sums, fused on
 * the GPU; rates land in mesh_dev[].prev_chem (zeroed first).  By default
 * one point-source pass; argot_hip_set_iterations(n) runs the reference's
 * transfer / smooth_photoion_rate / step_chemistry loop of at most n passes
 * and then update_chemistry_gpu.  One rank owning the whole box: a
 * this_run with nnode_* > 1, a nonzero rank or a local box other than the
 * global box gives ARGOT_ERR_ARG.  A ray whose source is so far outside the
 * box that it needs NSEG_PER_RAY or more segments gives
 * ARGOT_ERR_TOO_MANY_SEGMENTS (the reference asserts). */
void step_radiation_tree(struct fluid_mesh *mesh, struct radiation_src *src,
                         struct run_param *this_run, struct mpi_param *this_mpi,
                         struct cuda_mem_space *cuda_mem,
                         struct cuda_param *this_cuda, float dtime);

/* ------------------------------------------------------------------ */
/* ART diffuse-photon transfer (src/diffuse_photon/), one rank          */
/* ------------------------------------------------------------------ */
/* Layouts of src/diffuse_photon/diffuse_photon.h:77-160 with __HELIUM__
 * on, __HELIUM_BB__ off and __USE_GPU__ on. */
struct angle_info {                     /* 16 B, diffuse_photon.h:77-80 */
    float xovr, yovr, zovr;
    short base_id, corner_id;
};
struct step_func_factor {               /* 24 B, diffuse_photon.h:136-143 */
    float HI[2], HeI[2], HeII[2];       /* [0] ionization, [1] heating */
};
struct radiation_mesh;                  /* library-owned device accumulators */
struct omp_param;                       /* unused */
struct host_diffuse_param {             /* 48 B, diffuse_photon.h:146-153 */
    int corner_id_num[8];
    struct step_func_factor *step_fact;
    struct angle_info *angle;           /* [12*N_SIDE^2] */
};
struct cuda_diffuse_param {             /* 24 B, diffuse_photon.h:156-160 */
    struct step_func_factor *step_fact;
    struct angle_info *angle;
    struct radiation_mesh *rmesh;
};

/* replaces init_hd_param / free_hd_param, src/diffuse_photon/init_hd_param.c:8-30:
 * host arrays for N_ANG = 12*N_SIDE^2 directions (N_SIDE: see
 * argot_hip_set_diffuse_nside) */
void init_hd_param(struct host_diffuse_param *hd_param);
void free_hd_param(struct host_diffuse_param *hd_param);

/* replaces setup_diffuse_photon, src/diffuse_photon/setup_diffuse.c:8-14:
 * HEALPix directions (set_angle_info, sort_angle_id) and the step-function
 * factors (set_step_func_factor) */
void setup_diffuse_photon(struct host_diffuse_param *hd_param, struct run_param *this_run);

/* replaces init_diffuse_gpu, src/diffuse_photon/init_diffuse_gpu.cu:8-40,
 * and send_diffuse_data, src/diffuse_photon/send_diffuse_data.cu:8-35:
 * device buffers for the local mesh (even sizes, cubic cells), then the
 * directions and factors */
void init_diffuse_gpu(struct cuda_diffuse_param *cd_param, struct cuda_param *this_cuda,
                      struct run_param *this_run);
void send_diffuse_data(struct host_diffuse_param *hd_param, struct cuda_diffuse_param *cd_param,
                       struct cuda_param *this_cuda);

/* replaces calc_diffuse_photon_radiation (GPU variant),
 * src/diffuse_photon/calc_diffuse_photon_radiation.c:38-120 and
 * src/diffuse_photon/diffuse_prototype.h:30: absorption and recombination
 * emissivity per cell, every direction's parallel rays, per-cell rates
 * averaged over directions and added to mesh_dev[].prev_chem. */
void calc_diffuse_photon_radiation(struct fluid_mesh *mesh, struct run_param *this_run,
                                   struct cuda_mem_space *cuda_mem, struct cuda_param *this_cuda,
                                   struct host_diffuse_param *hd_param,
                                   struct cuda_diffuse_param *cd_param,
                                   struct omp_param *this_omp);

/* replaces free_diffuse_gpu, src/diffuse_photon/free_diffuse_gpu.cu:8-18 */
void free_diffuse_gpu(struct cuda_diffuse_param *cd_param, struct cuda_param *this_cuda);

/* ART controls (no reference counterpart): N_SIDE (reference: compile-time
 * 8, run_param.h:159; takes effect at init_hd_param/setup/init), the
 * __HEATCOOL__ step factors (default on, src/Makefile:187), directions
 * traced per batch (0, the default: as many as 8 GiB of accumulator slices
 * hold, at most 128; read by init_diffuse_gpu), Kahan-compensated sums over
 * directions. */
void argot_hip_set_diffuse_nside(int nside);
void argot_hip_set_diffuse_heatcool(int on);
void argot_hip_set_diffuse_batch(int ndir);
void argot_hip_set_diffuse_kahan(int on);
/* 1: exp and expm1 of the band optical depths in FP64 rounded to FP32, as
 * the reference's calc_intensity; 0 (default): FP32 exponential with a
 * Taylor series for 1 - exp(-tau) below 0.25. */
void argot_hip_set_diffuse_exp64(int on);
struct argot_diffuse_stats {
    uint64_t ndir;           /* directions */
    uint64_t ncross;         /* ray-cell crossings */
    double ms_trace;         /* ray tracing kernels only (not the slice fill), HIP events */
    double ms_gh;            /* per-direction rate kernels, HIP events */
    double ms_total;         /* whole call, wall clock */
};
void argot_hip_diffuse_stats(struct argot_diffuse_stats *out);

/* ------------------------------------------------------------------ */
/* Library controls (no reference counterpart)                         */
/* ------------------------------------------------------------------ */

enum argot_hip_status {
    ARGOT_OK = 0,
    ARGOT_ERR_HIP = 1,          /* a HIP runtime call failed */
    ARGOT_ERR_ARG = 2,          /* bad argument or size */
    ARGOT_ERR_STATE = 3,        /* init_gpu not called / mesh not sent /
                                   rate mode 1 without a trained surrogate */
    ARGOT_ERR_NOMEM = 4,        /* device or host allocation failed */
    ARGOT_ERR_TOO_MANY_SEGMENTS = 5, /* ray crosses too many domains */
    ARGOT_ERR_IO = 6            /* a file could not be opened, or is short */
};

/* Status of the last entry point called on this thread (0 = OK). */
int argot_hip_last_error(void);
/* For ARGOT_ERR_HIP / ARGOT_ERR_NOMEM from the step and transfer entry
 * points: the failing call's source line and HIP error string. */
const char *argot_hip_last_error_detail(void);
const char *argot_hip_error_string(int code);

/* n_He/n_H used for the helium number densities (reference: the
 * compile-time HELIUM_FACT, chemistry.h:25, default 0.0811258278). */
void argot_hip_set_helium_fact(double helium_fact);

/* Optical-depth summation: 0 = the reference's plain FP32 sums (default),
 * 1 = Kahan-compensated FP32 sums. */
void argot_hip_set_tau_mode(int mode);

/* Rate integral of step_radiation_tree: 0 = exact frequency integral
 * (default), 1 = MLP surrogate with exact fallback (see below): a
 * wavefront's visit of a tree node is served by the network only when every
 * emitting lane's ray qualifies, otherwise every lane takes the exact
 * integral, so per-cell rates depend on the target tile shape within the
 * validated tolerance. */
void argot_hip_set_rate_mode(int mode);

/* Counters of the last step_radiation_tree call on this thread (of the
 * context that call used). */
struct argot_step_stats {
    uint64_t ntarget;        /* target cells */
    uint64_t nray;           /* rays (== segments, one rank) */
    uint64_t ncross;         /* cell crossings incl. the end-cell term */
    uint64_t nfallback;      /* rays whose rates took the FP64 reference loop */
    double ms_build;         /* host tree build + upload */
    double ms_kernel;        /* fused kernel(s), HIP events */
    double ms_total;         /* whole call, wall clock */
    uint64_t nsurrogate;     /* rays whose rates came from the MLP surrogate */
    uint64_t iterations;     /* transfer passes (1 unless argot_hip_set_iterations) */
    double ms_chem;          /* smoothing + chemistry kernels over all passes */
    uint64_t nvisit;         /* tree nodes visited by the wavefronts (64 targets each) */
    uint64_t nemit_iter;     /* visits in which at least one lane emitted a ray:
                                nray / (64 nemit_iter) is the lane utilisation of
                                the walk and the rate loop */
    double ms_walk;          /* walk kernel of the last calc_optical_depth_cuda
                                call (segment list already on the device), HIP
                                events */
    double ms_rate;          /* rate kernel of the last calc_photoion_rate_cuda
                                call, HIP events */
    uint64_t nspec_call;     /* distinct spectra among that call's rays */
    uint64_t noverflow;      /* rays needing NSEG_PER_RAY or more segments (a
                                source far outside the box); nonzero makes the
                                step return ARGOT_ERR_TOO_MANY_SEGMENTS, where
                                the reference asserts (calc_ray_segment.c:212) */
};
void argot_hip_step_stats(struct argot_step_stats *out);

/* ------------------------------------------------------------------ */
/* Rate surrogate (no reference counterpart)                            */
/* ------------------------------------------------------------------ */
/* A 3 -> 64 -> 64 -> 6 SiLU network with bf16 weights maps the log10
 * threshold optical depths (HI, HeI, HeII) of a ray to log10 of the
 * unit-spectrum kernels F_s = sum_k S_k sigma_s(k) exp(-tau_k) and
 * G_s = sum_k S_k sigma_s(k) (nu_k - nu_s) exp(-tau_k), s = HI, HeI, HeII.
 * The photon-conserving cell average of a ray is the 2-point
 * Gauss-Legendre mean of the network along the in-cell optical-depth path.
 * It is trained on the GPU (bf16 MFMA forward and backward, FP32
 * accumulation, Adam) on samples of the exact integral over the given
 * frequency grid and spectrum shape, then validated against that integral
 * on a trust grid over its input hull: a ray is served by the network only
 * if its spectrum has the trained shape, its path lies in the hull and in
 * trusted cells, and its in-cell optical depth is within dtau_max;
 * otherwise the exact integral runs.
 *
 * Two sample distributions: uniform over the hull (nseed = 0), or the
 * optical depths a caller's rays occupy (nseed > 0: each sample is one of
 * the seed points plus Gaussian jitter, clamped to the hull), with
 * validation points drawn the same way and binned into the trust grid (a
 * cell is trusted when at least 8 validation points fall in it and all are
 * within trust_tol).  poly_degree 3 fits a cubic polynomial of the
 * normalised input to each output by least squares and trains the network
 * on the remainder. */
struct argot_surrogate_config {
    int ngrid;                 /* frequency bins of the integral */
    const double *nu;          /* [ngrid] nu / nu_L */
    const float *cs;           /* [3*ngrid] sigma HI, HeI, HeII per bin (cm^2) */
    const double *spec;        /* [ngrid] photon rate per bin (shape; any scale) */
    int64_t nsample;           /* training samples, multiple of batch */
    int batch;                 /* samples per optimizer step, multiple of 64 */
    int nstep;                 /* optimizer steps */
    float lr_max, lr_min;      /* Adam step size: 2% linear warm-up, cosine decay */
    uint64_t seed;             /* samples, initial weights and validation points */
    float log_tau_lo, log_tau_hi;  /* input hull per axis (log10 tau) */
    int trust_n;               /* trust cells per axis */
    float trust_tol;           /* max relative kernel error of a trusted cell */
    float dtau_max;            /* max in-cell optical depth for the 2-point rule */
    const float *seed_log_tau; /* [nseed*3] log10 threshold optical depths, or NULL */
    int64_t nseed;             /* seed points (0: uniform samples over the hull) */
    float jitter;              /* std of the Gaussian jitter around a seed (log10) */
    int poly_degree;           /* 0: none; 3: cubic baseline under the network */
    int64_t nvalidate;         /* validation points with seeds (0: 64 per trust cell) */
};
struct argot_surrogate_report {
    double loss_first;         /* normalised MSE of the first step's batch */
    double loss_last;          /* mean over the last 1% of steps */
    double ms_samples, ms_train, ms_validate;
    double err_max, err_p99, err_median;   /* relative kernel error over the validation points */
    double trusted_frac;       /* uniform: fraction of trust cells within trust_tol;
                                  seeds: fraction of validation points in trusted cells */
};
/* A log frequency grid for the surrogate: ngrid bins over
 * [10^log_nu_min, 10^log_nu_max] nu_L with centres nu[i] (the convention
 * of setup_freq_param), cross sections cs[3*i + s] (setup_cross_section's
 * fits) and a blackbody photon spectrum per bin (setup_photon_rate's
 * quadrature) normalised to a unit total. */
void argot_hip_spectral_grid(int ngrid, double log_nu_min, double log_nu_max, float temperature,
                             double *nu, float *cs, double *spec);
/* Generate samples, train, validate; the network then serves
 * step_radiation_tree when argot_hip_set_rate_mode(1) is set and the
 * step's frequency grid has ngrid bins. */
void argot_hip_surrogate_train(const struct argot_surrogate_config *cfg,
                               struct argot_surrogate_report *rep);
/* The first n training samples of cfg's generator (log10 tau [n*3],
 * log10 kernels [n*6]), without training. */
void argot_hip_surrogate_samples(const struct argot_surrogate_config *cfg, int64_t n,
                                 float *log_tau, float *log_kernel);
/* FP32 master parameters (ARGOT_SURROGATE_NPARAM floats: W1[3][64],
 * b1[64], W2[64][64], b2[64], W3[64][6], b3[6]) and normalisation
 * (xc, xs, ymean[6], ystd[6]: input (log10 tau - xc) * xs, output
 * log10 F = ymean + ystd * y). */
#define ARGOT_SURROGATE_NPARAM 4806
void argot_hip_surrogate_get_params(float *params, float *norm);
/* Sets the network (and drops any baseline; see below). */
void argot_hip_surrogate_set_params(const float *params, const float *norm);
/* Network outputs (log10 kernels [n*6]) for log10 tau [n*3]. */
void argot_hip_surrogate_eval(const float *log_tau, float *log_kernel, int64_t n);
/* Gradient of the normalised mean squared error over n samples (n a
 * multiple of 64) with respect to the parameters, as the training kernel
 * computes it (straight through the bf16 rounding of the weights). */
void argot_hip_surrogate_gradient(const float *log_tau, const float *log_kernel, int64_t n,
                                  float *grad, double *loss);
/* Validated relative error per trust cell (trust_n^3 floats, x slowest;
 * with seeds, cells without validation points read as infinity). */
void argot_hip_surrogate_trust_error(float *err, int64_t n);
/* The polynomial baseline under the network: ARGOT_SURROGATE_NPOLY
 * monomials of the normalised input z = (log10 tau - xc) * xs, in the order
 * 1, z0, z1, z2, z0z0, z0z1, z0z2, z1z1, z1z2, z2z2, z0z0z0, z0z0z1, z0z0z2,
 * z0z1z1, z0z1z2, z0z2z2, z1z1z1, z1z1z2, z1z2z2, z2z2z2, times coef[m*6 + o]
 * added to output o.  get returns the number of monomials in use (0 or
 * ARGOT_SURROGATE_NPOLY); set takes npoly 0 (no baseline) or
 * ARGOT_SURROGATE_NPOLY. */
#define ARGOT_SURROGATE_NPOLY 20
int argot_hip_surrogate_get_baseline(float *coef);
void argot_hip_surrogate_set_baseline(const float *coef, int npoly);

/* Opening angle for step_radiation_tree (default 0.7 = ARGOT_THETA_CRIT). */
void argot_hip_set_theta(float theta);

/* Persistent grid of the fused kernel, in wavefronts per compute unit
 * (default 24: six per SIMD, the exact-rate kernels' register budget;
 * n <= 0 restores the default). */
void argot_hip_set_blocks_per_cu(int n);

/* Target tile of one wavefront in the fused kernel: tx*ty*tz = 64 cells,
 * each a power of two (default 4x4x4). */
void argot_hip_set_tile(int tx, int ty, int tz);

/* 1 (default): step_radiation_tree replays its persistent fused launch
 * from a HIP graph, captured again only when a launch argument changes;
 * 0: plain stream launch. */
void argot_hip_set_graph(int on);
/* Graph captures so far (tests). */
int64_t argot_hip_graph_captures(void);

/* Autotuner: times the step on the caller's inputs for each tile shape and
 * persistent-grid size, keeps the fastest, and reports every candidate. */
#define ARGOT_TUNE_MAX 32
struct argot_tune_entry {
    int tile[3];
    int blocks_per_cu;
    double ms_kernel;
};
struct argot_tune_report {
    int n;                   /* candidates run */
    int best;                /* index of the selected one */
    struct argot_tune_entry entry[ARGOT_TUNE_MAX];
};
void argot_hip_autotune_step(struct fluid_mesh *mesh, struct radiation_src *src,
                             struct run_param *this_run, struct cuda_mem_space *cuda_mem,
                             struct cuda_param *this_cuda, struct argot_tune_report *rep);

/* Measured device copy bandwidth: best of reps copies of nbytes with a
 * plain streaming kernel, in GB/s of bytes read plus written. */
double argot_hip_copy_probe(int64_t nbytes, int reps);

#ifdef __cplusplus
}
#endif
#endif /* ARGOT_HIP_H */
