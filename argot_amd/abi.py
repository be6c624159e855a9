"""ctypes mirror of include/argot_hip.h (the C-ABI of libargot_hip.so).

The structs have the reference's byte layout (src/radiation.h, src/fluid.h,
src/run_param.h, src/source.h, src/cuda_mem_space.h; __HELIUM__ on,
__HYDROGEN_MOL__ off, NGRID_NU = 256).  tests/test_abi.py checks every size
and offset against the compiled library.
"""
import ctypes as C

import numpy as np

NGRID_NU = 256
NSEG_PER_RAY = 5
NMAX_DEV = 4


class PrimChem(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "wmol", "felec", "fHI", "fHII", "fHeI", "fHeII", "fHeIII",
        "GammaHI", "HeatHI", "GammaHeI", "GammaHeII", "HeatHeI", "HeatHeII")]


class FluidMesh(C.Structure):
    _fields_ = ([(n, C.c_float) for n in (
        "dens", "eneg", "momx", "momy", "momz", "uene", "duene", "durad", "etrp",
        "pot", "dens_prev")]
        + [("chem", PrimChem), ("prev_chem", PrimChem), ("prev_uene", C.c_float),
           ("under_shock", C.c_short), ("high_mach", C.c_short)])


class FreqParam(C.Structure):
    _fields_ = [("log_nu_min", C.c_double), ("log_nu_max", C.c_double),
                ("dlog_nu", C.c_double), ("nu", C.c_double * NGRID_NU)]


class RadiationSrc(C.Structure):
    _fields_ = [("type", C.c_int), ("param", C.c_float), ("xpos", C.c_float),
                ("ypos", C.c_float), ("zpos", C.c_float),
                ("photon_rate", C.c_double * NGRID_NU)]


class CrossSection(C.Structure):
    _fields_ = [("csect_HI", C.c_float), ("csect_HeI", C.c_float), ("csect_HeII", C.c_float)]


class Cosmology(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("omega_m", "omega_v", "omega_b", "omega_nu",
                                         "hubble", "tend")]


class RunParam(C.Structure):
    _fields_ = [
        ("model_name", C.c_void_p), ("proc_file", C.c_void_p), ("diag_file", C.c_void_p),
        ("err_file", C.c_void_p),
        ("helium_flag", C.c_int), ("hydrogen_mol_flag", C.c_int), ("cosmology_flag", C.c_int),
        ("ngrid_nu", C.c_int), ("mpi_rank", C.c_int), ("mpi_nproc", C.c_int),
        ("rank_x", C.c_int), ("rank_y", C.c_int), ("rank_z", C.c_int),
        ("nnode_x", C.c_int), ("nnode_y", C.c_int), ("nnode_z", C.c_int),
        ("lunit", C.c_double), ("munit", C.c_double), ("tunit", C.c_double),
        ("eunit", C.c_double), ("denstonh", C.c_double), ("uenetok", C.c_double),
        ("masstonh", C.c_double),
        ("anow", C.c_float), ("tnow", C.c_float), ("znow", C.c_float), ("hnow", C.c_float),
        ("cosm", Cosmology), ("tend", C.c_float), ("dtime", C.c_float),
        ("freq", FreqParam), ("csect", CrossSection * NGRID_NU),
        ("nspecies", C.c_int), ("nchannel", C.c_int), ("step", C.c_int),
        ("output_indx", C.c_int), ("noutput", C.c_int), ("output_timing", C.c_void_p),
        ("nsrc", C.c_uint64), ("nray", C.c_uint64), ("nseg", C.c_uint64),
        ("xmin", C.c_float), ("ymin", C.c_float), ("zmin", C.c_float),
        ("xmax", C.c_float), ("ymax", C.c_float), ("zmax", C.c_float),
        ("xmin_local", C.c_float), ("ymin_local", C.c_float), ("zmin_local", C.c_float),
        ("xmax_local", C.c_float), ("ymax_local", C.c_float), ("zmax_local", C.c_float),
        ("nmesh_x_total", C.c_int), ("nmesh_y_total", C.c_int), ("nmesh_z_total", C.c_int),
        ("nmesh_x_local", C.c_int), ("nmesh_y_local", C.c_int), ("nmesh_z_local", C.c_int),
        ("delta_x", C.c_float), ("delta_y", C.c_float), ("delta_z", C.c_float),
        ("theta_crit", C.c_float), ("nmesh_per_loop", C.c_int),
    ]


class PhotoionRate(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("GammaHI", "HeatHI", "GammaHeI", "GammaHeII",
                                         "HeatHeI", "HeatHeII")]


class RaySegment(C.Structure):
    _fields_ = [("ray_indx", C.c_uint64), ("local_rank", C.c_int), ("target_rank", C.c_int),
                ("xpos_start", C.c_float), ("ypos_start", C.c_float), ("zpos_start", C.c_float),
                ("xpos_end", C.c_float), ("ypos_end", C.c_float), ("zpos_end", C.c_float),
                ("optical_depth_HI", C.c_float), ("optical_depth_HeI", C.c_float),
                ("optical_depth_HeII", C.c_float)]


class LightRay(C.Structure):
    _fields_ = [("src", RadiationSrc), ("num_segment", C.c_int),
                ("segment", RaySegment * NSEG_PER_RAY),
                ("ix_target", C.c_int), ("iy_target", C.c_int), ("iz_target", C.c_int),
                ("optical_depth_HI", C.c_float), ("optical_depth_HeI", C.c_float),
                ("optical_depth_HeII", C.c_float)]


class LightRayIO(C.Structure):
    _fields_ = [("src", RadiationSrc),
                ("ix_target", C.c_int), ("iy_target", C.c_int), ("iz_target", C.c_int),
                ("optical_depth_HI", C.c_float), ("optical_depth_HeI", C.c_float),
                ("optical_depth_HeII", C.c_float)]


class CudaParam(C.Structure):
    _fields_ = [("num_cuda_dev", C.c_int), ("nmesh_per_dev", C.c_int), ("cuda_nblock", C.c_int),
                ("max_thread_dimx", C.c_int), ("max_block_dimx", C.c_int),
                ("strm", C.c_void_p * NMAX_DEV), ("diffuse_strm", C.c_void_p * NMAX_DEV)]


class CudaMemSpace(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "mesh_dev", "this_run_dev", "freq_dev", "segment_dev", "rate_dev", "chem_updated",
        "ray_IO_dev", "optical_depth_dev", "dtime_dev", "diff_chem_dev", "diff_uene_dev")]


class StepStats(C.Structure):
    _fields_ = [("ntarget", C.c_uint64), ("nray", C.c_uint64), ("ncross", C.c_uint64),
                ("nfallback", C.c_uint64), ("ms_build", C.c_double), ("ms_kernel", C.c_double),
                ("ms_total", C.c_double), ("nsurrogate", C.c_uint64),
                ("iterations", C.c_uint64), ("ms_chem", C.c_double),
                ("nvisit", C.c_uint64), ("nemit_iter", C.c_uint64), ("ms_walk", C.c_double),
                ("ms_rate", C.c_double), ("nspec_call", C.c_uint64),
                ("noverflow", C.c_uint64)]


class ChemStats(C.Structure):
    _fields_ = [("ncell", C.c_uint64), ("nsub", C.c_uint64), ("niter", C.c_uint64),
                ("ms_kernel", C.c_double), ("nstopped", C.c_uint64)]


class SurrogateConfig(C.Structure):
    _fields_ = [("ngrid", C.c_int), ("nu", C.POINTER(C.c_double)), ("cs", C.POINTER(C.c_float)),
                ("spec", C.POINTER(C.c_double)), ("nsample", C.c_int64), ("batch", C.c_int),
                ("nstep", C.c_int), ("lr_max", C.c_float), ("lr_min", C.c_float),
                ("seed", C.c_uint64), ("log_tau_lo", C.c_float), ("log_tau_hi", C.c_float),
                ("trust_n", C.c_int), ("trust_tol", C.c_float), ("dtau_max", C.c_float),
                ("seed_log_tau", C.POINTER(C.c_float)), ("nseed", C.c_int64),
                ("jitter", C.c_float), ("poly_degree", C.c_int), ("nvalidate", C.c_int64)]


class SurrogateReport(C.Structure):
    _fields_ = [("loss_first", C.c_double), ("loss_last", C.c_double),
                ("ms_samples", C.c_double), ("ms_train", C.c_double), ("ms_validate", C.c_double),
                ("err_max", C.c_double), ("err_p99", C.c_double), ("err_median", C.c_double),
                ("trusted_frac", C.c_double)]


SURROGATE_NPARAM = 4806
SURROGATE_NPOLY = 20


# One captured ray (test mode of step_radiation_tree), 120 bytes.
RAY_RECORD = np.dtype([
    ("target", np.int64), ("seq", np.int32), ("kind", np.int32), ("id", np.int32),
    ("ncross", np.int32), ("start", np.float32, 3), ("src", np.float32, 3),
    ("tau", np.float32, 3), ("surrogate", np.float32), ("hash", np.uint64), ("rate", np.float64, 6)])


def layout_expected():
    """Sizes/offsets in the order argot_hip_abi_layout() returns them."""
    return [
        C.sizeof(PrimChem), C.sizeof(FluidMesh), C.sizeof(FreqParam), C.sizeof(RadiationSrc),
        C.sizeof(CrossSection), C.sizeof(RunParam), C.sizeof(PhotoionRate),
        C.sizeof(RaySegment), C.sizeof(LightRay), C.sizeof(LightRayIO), C.sizeof(CudaParam),
        C.sizeof(CudaMemSpace),
        RunParam.lunit.offset, RunParam.denstonh.offset, RunParam.freq.offset,
        RunParam.csect.offset, RunParam.nray.offset, RunParam.nseg.offset,
        RunParam.xmin_local.offset, RunParam.delta_x.offset,
        FluidMesh.chem.offset, FluidMesh.prev_chem.offset, LightRayIO.optical_depth_HI.offset,
        RAY_RECORD.itemsize,
    ]


def mesh_array(n_cells):
    """Zeroed host array of struct fluid_mesh, viewable as numpy."""
    return (FluidMesh * n_cells)()


def mesh_view(arr):
    """numpy structured view (no copy) of a ctypes FluidMesh array."""
    return np.ctypeslib.as_array(arr)


# ART diffuse photons (include/argot_hip.h; __HELIUM__ on, __HELIUM_BB__ off)
class AngleInfo(C.Structure):
    _fields_ = [("xovr", C.c_float), ("yovr", C.c_float), ("zovr", C.c_float),
                ("base_id", C.c_short), ("corner_id", C.c_short)]


class StepFuncFactor(C.Structure):
    _fields_ = [("HI", C.c_float * 2), ("HeI", C.c_float * 2), ("HeII", C.c_float * 2)]


class HostDiffuseParam(C.Structure):
    _fields_ = [("corner_id_num", C.c_int * 8), ("step_fact", C.POINTER(StepFuncFactor)),
                ("angle", C.POINTER(AngleInfo))]


class CudaDiffuseParam(C.Structure):
    _fields_ = [("step_fact", C.c_void_p), ("angle", C.c_void_p), ("rmesh", C.c_void_p)]


class DiffuseStats(C.Structure):
    _fields_ = [("ndir", C.c_uint64), ("ncross", C.c_uint64), ("ms_trace", C.c_double),
                ("ms_gh", C.c_double), ("ms_total", C.c_double)]


class TuneEntry(C.Structure):
    _fields_ = [("tile", C.c_int * 3), ("blocks_per_cu", C.c_int), ("ms_kernel", C.c_double)]


TUNE_MAX = 32


class TuneReport(C.Structure):
    _fields_ = [("n", C.c_int), ("best", C.c_int), ("entry", TuneEntry * TUNE_MAX)]


_NP_SCALAR = {C.c_uint64: np.uint64, C.c_int64: np.int64, C.c_int: np.int32, C.c_uint: np.uint32,
              C.c_float: np.float32, C.c_double: np.float64}


def flat_dtype(ctype):
    """numpy dtype with the field offsets and size of a ctypes structure of
    scalar fields (ctypes' own buffer export guesses the tail padding)."""
    names = [f for f, _ in ctype._fields_]
    return np.dtype({"names": names,
                     "formats": [_NP_SCALAR[t] for _, t in ctype._fields_],
                     "offsets": [getattr(ctype, f).offset for f in names],
                     "itemsize": C.sizeof(ctype)})


def view_at(ctype, address, count):
    """numpy structured view (no copy) of count structures at address."""
    buf = (C.c_char * (C.sizeof(ctype) * count)).from_address(address)
    return np.frombuffer(buf, dtype=flat_dtype(ctype))
