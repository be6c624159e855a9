/*
 * ORACLE — test infrastructure, not product code.
 *
 * A plain-C CPU restatement of ARGOT's point-source radiation step, used
 * only as the checker: by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py.  Nothing in argot_amd/ links, loads or calls
 * it.  Every function cites the reference file:line it restates.
 *
 * Numerical contract: built with -O2 -fno-fast-math -ffp-contract=off, so
 * every float/double operation rounds exactly where the C source says it
 * does (the reference's own op order, its float/double promotions, IEEE
 * division and square root).  That is the "reference FP32 path" the HIP
 * kernels are held bit-exact against.  The *_exact functions are the FP64
 * oracle: exact slab intersections, expm1 for 1-exp(-x), FP64 sums.
 *
 * Conventions (reference run_param.h / fluid.h):
 *   cell (ix,iy,iz) of an nx*ny*nz mesh lives at iz + nz*(iy + ny*ix);
 *   one rank (NNODE = 1): local box == global box;
 *   three species HI, HeI, HeII (__HELIUM__ on, __HYDROGEN_MOL__ off).
 */
#ifndef ARGOT_ORACLE_H
#define ARGOT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Frequency grid, spectra, cross sections                             */
/* ------------------------------------------------------------------ */

/* setup_freq_param, src/source.c:38-53 (H2 off: log nu in [0, 1.5]). */
void orc_freq_grid(int ngrid, double log_nu_min, double log_nu_max,
                   double *nu, double *dlog_nu);

/* setup_photon_rate for a blackbody, src/source.c:10-21,56-138. */
void orc_photon_rate_bb(int ngrid, double log_nu_min, double log_nu_max,
                        float temperature, double total_rate, double *out);

/* csectHI / csectHeI / csectHeII, src/prim_chem/cross_section.c:20-53. */
double orc_csect_HI(double nu);
double orc_csect_HeI(double nu);
double orc_csect_HeII(double nu);

/* setup_cross_section, src/cross_section.c:5-22: cs[3*i + s], FP32. */
void orc_cross_sections(int ngrid, const double *nu, float *cs);

/* ------------------------------------------------------------------ */
/* Problem description                                                 */
/* ------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nz;                      /* local mesh (== global, one rank) */
    float xmin, ymin, zmin;              /* box corner */
    float xmax, ymax, zmax;
    float dx, dy, dz;                    /* cell sizes (run_param delta_x..) */
    double lunit, denstonh;              /* cm per box unit, nH per code dens */
    double helium_fact;                  /* nHe/nH (chemistry.h:25) */
    float theta_crit;                    /* opening angle (run_param.h:80) */
    int ngrid;                           /* frequency bins */
    const double *nu;                    /* [ngrid] in units of nu_L */
    const float *cs;                     /* [ngrid*3] HI, HeI, HeII */
} orc_grid;

/* Mesh fields (FP32, z fastest). */
typedef struct {
    const float *dens, *fHI, *fHeI, *fHeII;
} orc_fields;

/* ------------------------------------------------------------------ */
/* Source oct-tree (setup_light_ray_tree.c)                            */
/* ------------------------------------------------------------------ */

typedef struct orc_tree orc_tree;

/* construct_tree + init_tree_param, src/setup_light_ray_tree.c:465-585.
 * pos[3*i], spec[i*ngrid + k]: the nsrc radiation_src records. */
orc_tree *orc_tree_build(const orc_grid *g, int nsrc, const float *pos,
                         const double *spec);
void orc_tree_free(orc_tree *t);
int orc_tree_nnode(const orc_tree *t);

/* Node record access (for tests): cm, length, n, spectrum row. */
void orc_tree_node(const orc_tree *t, int node, float *cm, float *length,
                   int *n, int *nchild, const double **spec);
/* Sources below node k (original indices, sorted-key order); returns n. */
int orc_tree_node_members(const orc_tree *t, int node, int *out);

/* count_ray_to, src/setup_light_ray_tree.c:587-605 + count_src :241-283 */
uint64_t orc_count_ray_to(const orc_tree *t, const orc_grid *g,
                          int ix, int iy, int iz);

/* setup_light_ray_to, src/setup_light_ray_tree.c:321-395.
 * Writes up to cap rays in the reference's order.  For ray r:
 *   kind[r] = 0 (tree node, id = node index) or 1 (source, id = source
 *   index); pos[3r] = its position (node cm or source position).
 * Returns the number of rays (may exceed cap; extra rays not written). */
uint64_t orc_rays_to(const orc_tree *t, const orc_grid *g, int ix, int iy,
                     int iz, uint64_t cap, int *kind, int *id, float *pos);

/* Spectrum of a ray source: kind 0 -> node, 1 -> source. */
const double *orc_ray_spec(const orc_tree *t, int kind, int id);

/* ------------------------------------------------------------------ */
/* Ray segments and the optical-depth walk                             */
/* ------------------------------------------------------------------ */

/* calc_ray_segment, src/calc_ray_segment.c:15-213, one rank (NNODE=1).
 * From source position and target cell to segment start/end (global
 * coordinates).  Returns the number of segments written (<= maxseg). */
int orc_ray_segment(const orc_grid *g, const float *src, int ix, int iy,
                    int iz, int maxseg, float *start, float *end);

/* calc_optical_depth, src/calc_optical_depth.c:20-228, one segment.
 * tau[3] out (HI, HeI, HeII).  If cells != NULL the visited cell indices
 * (loop cells, then the end cell) are written up to cap.
 * Returns N_cross = loop iterations + 1 (the end cell term). */
int orc_walk_fp32(const orc_grid *g, const orc_fields *f, const float *start,
                  const float *end, float *tau, int cap, int *cells);

/* Same walk with Kahan-compensated FP32 sums of the per-cell terms. */
int orc_walk_fp32_kahan(const orc_grid *g, const orc_fields *f,
                        const float *start, const float *end, float *tau);

/* FP64 oracle: exact chord lengths through the cells the straight segment
 * crosses, tau_s = sum n_s * l in double.  n_s uses the same FP32 n_s as the
 * walk (dens*f) promoted to double.  Returns the number of cells crossed. */
int orc_walk_exact(const orc_grid *g, const orc_fields *f, const float *start,
                   const float *end, double *tau);

/* ------------------------------------------------------------------ */
/* Photoionization / photoheating rates                                */
/* ------------------------------------------------------------------ */

/* delta_length, src/calc_photoion_rate.c:52-90. */
float orc_delta_length(const orc_grid *g, float dx, float dy, float dz);

/* Per-ray photon-conserving rates, src/calc_photoion_rate.c:159-387.
 * tau_ray: the ray's FP32 optical depths in code units (accum result).
 * out[6] = GammaHI, HeatHI, GammaHeI, GammaHeII, HeatHeI, HeatHeII with the
 * heats already times hplanck*nuL (the merge scaling, :403-412).
 * mode 0: the reference's arithmetic (naive 1-exp(-dtau) in FP64);
 * mode 1: FP64 oracle (1-exp(-dtau) via -expm1). */
void orc_ray_rate(const orc_grid *g, const orc_fields *f, const float *src,
                  const double *spec, int ix, int iy, int iz,
                  const double *tau_ray, int mode, double *out);

/* ------------------------------------------------------------------ */
/* Whole point-source step (calc_ARGOT_part for one rank)              */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t nray;           /* sum over targets of count_ray_to */
    uint64_t ncross;         /* sum of N_cross over segments */
    uint64_t nseg;           /* segments walked */
} orc_stats;

/* Targets [t0, t1) in cell order.  rates_ref[6*N] receives the reference
 * arithmetic with its FP32 accumulation (prev_chem += per ray, ray order);
 * rates_f64[6*N] (may be NULL) the FP64 oracle (exact walk, expm1, FP64
 * sums).  Rates are written for targets in range only (others untouched).
 * nthreads <= 0: OpenMP default.  kahan != 0 uses the Kahan walk for the
 * reference-arithmetic result. */
void orc_step(const orc_grid *g, const orc_fields *f, const orc_tree *t,
              int64_t t0, int64_t t1, float *rates_ref, double *rates_f64,
              int kahan, int nthreads, orc_stats *st);

/* Per-ray records for targets [t0, t1) (for kernel parity tests):
 * target flat index, kind, id, tau (FP32 walk), N_cross and a 64-bit hash
 * of the visited cell sequence (h = (h ^ cell) * 0x100000001b3, FNV-1a
 * style, seeded with 0xcbf29ce484222325).  Returns rays written. */
uint64_t orc_ray_records(const orc_grid *g, const orc_fields *f,
                         const orc_tree *t, int64_t t0, int64_t t1,
                         uint64_t cap, int64_t *target, int *kind, int *id,
                         float *tau, int *ncross, uint64_t *hash);

/* Optical-depth-to-rate frequency integral in the surrogate's units:
 * F_s(tau) = sum_nu N_nu sigma_s(nu) exp(-sum_k tau_k sigma_k/sigma_k,th)
 * for a unit-normalised spectrum (used by surrogate tests). */
void orc_thin_kernel(int ngrid, const double *nu, const float *cs,
                     const double *spec, const double *tau_th, double *out4);

/* smooth_photoion_rate on one rank (src/smooth_photoion_rate.c:1058-1090):
 * 27-point stencil, central weight 30, ghost cells = nearest interior
 * cell; six rates per cell in mesh order. */
void orc_smooth_rates(int nx, int ny, int nz, const float *in, float *out);

/* ------------------------------------------------------------------ */
/* ART diffuse-photon transfer (art.c), one rank                        */
/* ------------------------------------------------------------------ */

/* chealpix pix2ang_ring from the published ring-scheme formulas. */
void orc_pix2ang_ring(long nside, long ipix, double *theta, double *phi);

/* set_angle_info (+ sort_angle_id when sorted != 0): ovr[3*i] direction
 * cosines (FP32), base_id, corner_id per direction; corner_num[8]. */
void orc_art_angles(int nside, int sorted, float *ovr, short *base, short *corner,
                    int *corner_num);

/* set_step_func_factor: f = {HI[0], HI[1], HeI[0], HeI[1], HeII[0], HeII[1]};
 * heatcool selects the __HEATCOOL__ 64-bin form. */
void orc_art_step_factors(int heatcool, float f[6]);

/* calc_rmesh_data: chem[7*c] = felec, fHI, fHII, fHeI, fHeII, fHeIII,
 * prev_uene; kap/src [3*c] for the three bands. */
/* The recombination fits of src/diffuse_photon/diffuse_chemistry.c:13-97
 * at temperature t: k02_A, k02_B, k04_A, k04_B, k06_A, k06_B. */
void orc_art_recomb(double t, double *out);
void orc_art_rmesh(long ncell, const float *dens, const float *chem, double denstonh,
                   double uenetok, double helium_fact, float *kap, float *src);

/* One direction (zeroed accumulators, rays in group order): per-cell path
 * length [cm] and band intensities I[3*c].  Returns the cell crossings. */
long orc_art_direction(const orc_grid *g, const float ovr[3], int base, const float *kap,
                       const float *src, float *length, float *I);

/* calc_GH_tot: gh[6*c] += per-direction rates (GHI, HHI, GHeI, HHeI,
 * GHeII, HHeII). */
void orc_art_gh_tot(long ncell, const float *length, const float *I, const float *kap,
                    const float *src, const float f[6], float *gh);

/* The whole pass (sorted directions, calc_GH_sum): rates[6*c] in mesh
 * order (GammaHI, HeatHI, GammaHeI, GammaHeII, HeatHeI, HeatHeII) +=.
 * Returns the total cell crossings. */
long orc_art_run(const orc_grid *g, int nside, int heatcool, const float *dens,
                 const float *chem, double uenetok, float *rates);

/* ---- chemistry (chem.c) ---- */

/* struct prim_chem (13 floats). */
typedef struct {
    float wmol, felec, fHI, fHII, fHeI, fHeII, fHeIII;
    float GammaHI, HeatHI, GammaHeI, GammaHeII, HeatHeI, HeatHeII;
} orc_chem;

/* Units and cosmology of struct run_param, plus the helium number fraction. */
typedef struct {
    double lunit, munit, tunit, uenetok, denstonh, helium_fact;
    float anow, znow;
} orc_chem_params;

/* k01..k06 (case B) at temperature t [K]. */
void orc_chem_k(double t, double *k);
/* creh2 crehe2 crexihe2 crehe3 cioh1 ciohe1 ciohe2 cexh1 cexhe1 cexhe2 brems compt. */
void orc_chem_cooling(double t, float zred, double *o);
/* calc_heatcool_rate_dev: heat - cool [erg cm^-3 s^-1]. */
double orc_chem_heatcool(const orc_chem *c, float zred, double nH, double T, double helium_fact);
/* advance_heatcool_dev: *uene over dtime (code units); returns the bisections. */
int orc_chem_advance_heatcool(float dens, float *uene, const orc_chem *c,
                              const orc_chem_params *p, float dtime);
/* calc_dtime_dev: safe step [s] (mesh_uene: the mesh's current energy). */
float orc_chem_dtime(float dens, float mesh_uene, const orc_chem *c, const orc_chem_params *p);
/* advance_reaction_and_heatcool_dev on one cell. */
void orc_chem_advance(float dens, float mesh_uene, float *uene, orc_chem *c,
                      const orc_chem_params *p, float dtime, int *nsub, int *niter);
/* The per-cell kernels of chemistry.cu over an AoS struct fluid_mesh array:
 * the step (maxima of the relative changes, work counts), the minimum safe
 * step over cells, and update_chemistry. */
void orc_chem_step_mesh(unsigned char *mesh, int64_t ncell, const orc_chem_params *p,
                        float dtime, float *diff_chem, float *diff_uene, uint64_t *nsub,
                        uint64_t *niter);
float orc_chem_dtime_mesh(const unsigned char *mesh, int64_t ncell, const orc_chem_params *p);
void orc_chem_update_mesh(unsigned char *mesh, int64_t ncell);

#ifdef __cplusplus
}
#endif
#endif
