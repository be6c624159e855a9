"""Loader for libargot_hip.so (built in-tree by `make -C argot_amd/csrc`).

There is no fallback: if the library is missing, or a GPU entry point fails,
the calls raise.
"""
import ctypes as C
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# ARGOT_HIP_LIB selects another build of the same library (e.g. a tuning
# variant); default: the in-tree build.
LIB_PATH = os.environ.get("ARGOT_HIP_LIB", os.path.join(HERE, "libargot_hip.so"))

# Entry points that include/argot_hip.h declares (checked by tests).
EXPORTS = [
    "setup_freq_param", "setup_photon_rate", "setup_cross_section",
    "init_gpu", "send_mesh_data", "recv_mesh_data", "send_run_param_data",
    "allocate_pinned_segment", "deallocate_pinned_segment",
    "allocate_pinned_light_ray_IO", "deallocate_pinned_light_ray_IO",
    "calc_optical_depth_cuda", "argot_hip_calc_optical_depth_noalloc",
    "calc_photoion_rate_cuda", "zero_out_photoion_rate_cuda",
    "merge_photoion_data", "recv_photoion_data", "send_photoion_data", "step_radiation_tree",
    "argot_hip_last_error", "argot_hip_error_string", "argot_hip_set_helium_fact",
    "argot_hip_set_tau_mode", "argot_hip_set_rate_mode", "argot_hip_step_stats",
    "argot_hip_set_theta", "argot_hip_set_blocks_per_cu", "argot_hip_copy_probe",
    "argot_hip_spectral_grid", "argot_hip_surrogate_train", "argot_hip_surrogate_samples",
    "argot_hip_surrogate_get_params", "argot_hip_surrogate_set_params",
    "argot_hip_surrogate_eval", "argot_hip_surrogate_gradient",
    "argot_hip_surrogate_trust_error", "argot_hip_surrogate_get_baseline",
    "argot_hip_surrogate_set_baseline",
    "init_hd_param", "free_hd_param", "setup_diffuse_photon", "init_diffuse_gpu",
    "send_diffuse_data", "calc_diffuse_photon_radiation", "free_diffuse_gpu",
    "argot_hip_set_diffuse_nside", "argot_hip_set_diffuse_heatcool",
    "argot_hip_set_diffuse_batch", "argot_hip_set_diffuse_kahan", "argot_hip_diffuse_stats",
    "argot_hip_set_diffuse_exp64", "argot_hip_set_tile", "argot_hip_set_graph",
    "argot_hip_graph_captures", "argot_hip_autotune_step", "argot_hip_last_error_detail",
    "smooth_photoion_rate", "step_chemistry_gpu", "step_chemistry", "calc_dtime_gpu",
    "calc_timestep_chem", "update_chemistry_gpu", "argot_hip_chem_stats",
    "argot_hip_set_iterations", "input_mesh_single", "input_mesh", "input_src_file", "input_src",
    "input_data", "output_mesh_single", "output_mesh", "output_src", "output_data",
    "argot_hip_set_max_sources", "argot_hip_set_chem_queue", "argot_hip_set_wide_index",
    "argot_hip_set_host_threads",
    "argot_hip_release_gpu",
]

_lib = None


class ArgotError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ArgotError(f"{LIB_PATH} is not built (run `make -C argot_amd/csrc`)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.setup_freq_param.argtypes = [P(abi.FreqParam)]
    L.setup_photon_rate.argtypes = [P(abi.FreqParam), P(abi.RadiationSrc), C.c_double]
    L.setup_cross_section.argtypes = [P(abi.CrossSection), P(abi.FreqParam)]
    cm, cp, rp = P(abi.CudaMemSpace), P(abi.CudaParam), P(abi.RunParam)
    L.init_gpu.argtypes = [C.c_void_p, cm, cp, rp]
    L.argot_hip_release_gpu.argtypes = [cm]
    L.argot_hip_release_gpu.restype = C.c_int
    L.send_mesh_data.argtypes = [C.c_void_p, cm, cp, rp]
    L.recv_mesh_data.argtypes = [C.c_void_p, cm, cp, rp]
    L.send_run_param_data.argtypes = [rp, cm, cp]
    L.allocate_pinned_segment.argtypes = [P(C.c_void_p), C.c_uint64]
    L.deallocate_pinned_segment.argtypes = [C.c_void_p]
    L.allocate_pinned_light_ray_IO.argtypes = [P(C.c_void_p), C.c_uint64]
    L.deallocate_pinned_light_ray_IO.argtypes = [C.c_void_p]
    L.calc_optical_depth_cuda.argtypes = [C.c_void_p, cm, cp, rp]
    L.argot_hip_calc_optical_depth_noalloc.argtypes = [C.c_void_p, cm, cp, rp]
    L.calc_photoion_rate_cuda.argtypes = [cm, C.c_void_p, C.c_void_p, cp, rp]
    L.zero_out_photoion_rate_cuda.argtypes = [cm, cp, rp]
    L.merge_photoion_data.argtypes = [C.c_void_p, cm, cp, rp]
    L.recv_photoion_data.argtypes = [C.c_void_p, cm, cp, rp]
    L.send_photoion_data.argtypes = [C.c_void_p, cm, cp, rp]
    L.step_radiation_tree.argtypes = [C.c_void_p, C.c_void_p, rp, C.c_void_p, cm, cp, C.c_float]
    L.argot_hip_last_error.restype = C.c_int
    L.argot_hip_error_string.argtypes = [C.c_int]
    L.argot_hip_error_string.restype = C.c_char_p
    L.argot_hip_last_error_detail.restype = C.c_char_p
    L.argot_hip_set_helium_fact.argtypes = [C.c_double]
    L.argot_hip_set_tau_mode.argtypes = [C.c_int]
    L.argot_hip_set_rate_mode.argtypes = [C.c_int]
    L.argot_hip_step_stats.argtypes = [P(abi.StepStats)]
    L.argot_hip_set_theta.argtypes = [C.c_float]
    L.argot_hip_set_blocks_per_cu.argtypes = [C.c_int]
    L.argot_hip_set_capture.argtypes = [C.c_void_p, C.c_int64]
    L.argot_hip_copy_probe.argtypes = [C.c_int64, C.c_int]
    L.argot_hip_copy_probe.restype = C.c_double
    L.argot_hip_record_size.restype = C.c_int64
    L.argot_hip_abi_layout.argtypes = [P(C.c_int64), C.c_int]
    L.argot_hip_abi_layout.restype = C.c_int
    L.argot_hip_segment_start.argtypes = [rp, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                          C.c_void_p]
    L.argot_hip_segment_start.restype = C.c_int
    L.argot_hip_tree_build.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                       C.c_float]
    L.argot_hip_tree_build.restype = C.c_void_p
    L.argot_hip_tree_free.argtypes = [C.c_void_p]
    L.argot_hip_tree_nnode.argtypes = [C.c_void_p]
    L.argot_hip_tree_rays_to.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
    L.argot_hip_tree_rays_to.restype = C.c_uint64
    L.argot_hip_tree_spec.argtypes = [C.c_void_p, C.c_int]
    L.argot_hip_tree_spec.restype = P(C.c_double)
    L.argot_hip_spectral_grid.argtypes = [C.c_int, C.c_double, C.c_double, C.c_float,
                                          C.c_void_p, C.c_void_p, C.c_void_p]
    L.argot_hip_surrogate_train.argtypes = [P(abi.SurrogateConfig), P(abi.SurrogateReport)]
    L.argot_hip_surrogate_samples.argtypes = [P(abi.SurrogateConfig), C.c_int64, C.c_void_p,
                                              C.c_void_p]
    L.argot_hip_surrogate_get_params.argtypes = [C.c_void_p, C.c_void_p]
    L.argot_hip_surrogate_set_params.argtypes = [C.c_void_p, C.c_void_p]
    L.argot_hip_surrogate_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    L.argot_hip_surrogate_gradient.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                               C.c_void_p]
    L.argot_hip_surrogate_trust_error.argtypes = [C.c_void_p, C.c_int64]
    L.argot_hip_surrogate_get_baseline.argtypes = [C.c_void_p]
    L.argot_hip_surrogate_get_baseline.restype = C.c_int
    L.argot_hip_surrogate_set_baseline.argtypes = [C.c_void_p, C.c_int]
    hd, cd = P(abi.HostDiffuseParam), C.c_void_p
    L.init_hd_param.argtypes = [hd]
    L.free_hd_param.argtypes = [hd]
    L.setup_diffuse_photon.argtypes = [hd, rp]
    L.init_diffuse_gpu.argtypes = [cd, cp, rp]
    L.send_diffuse_data.argtypes = [hd, cd, cp]
    L.calc_diffuse_photon_radiation.argtypes = [C.c_void_p, rp, cm, cp, hd, cd, C.c_void_p]
    L.free_diffuse_gpu.argtypes = [cd, cp]
    for n in ("argot_hip_set_diffuse_nside", "argot_hip_set_diffuse_heatcool",
              "argot_hip_set_diffuse_batch", "argot_hip_set_diffuse_kahan",
              "argot_hip_set_diffuse_exp64"):
        getattr(L, n).argtypes = [C.c_int]
    L.argot_hip_diffuse_stats.argtypes = [P(abi.DiffuseStats)]
    L.smooth_photoion_rate.argtypes = [C.c_void_p, rp, cm, cp, C.c_void_p]
    fp = P(C.c_float)
    L.step_chemistry_gpu.argtypes = [cm, cp, rp, fp, fp, C.c_float]
    L.step_chemistry.argtypes = [cm, cp, rp, C.c_float, fp, fp]
    L.calc_dtime_gpu.argtypes = [cm, cp, rp]
    L.calc_dtime_gpu.restype = C.c_double
    L.calc_timestep_chem.argtypes = [cm, cp, rp]
    L.calc_timestep_chem.restype = C.c_double
    L.update_chemistry_gpu.argtypes = [cm, cp, rp]
    L.argot_hip_chem_stats.argtypes = [P(abi.ChemStats)]
    L.argot_hip_set_iterations.argtypes = [C.c_int]
    L.argot_hip_set_chem_queue.argtypes = [C.c_int]
    L.argot_hip_set_wide_index.argtypes = [C.c_int]
    L.argot_hip_set_host_threads.argtypes = [C.c_int]
    vp, cs_ = C.c_void_p, C.c_char_p
    for n in ("input_mesh_single", "input_mesh", "output_mesh_single", "output_mesh"):
        getattr(L, n).argtypes = [vp, rp, cs_]
    for n in ("input_src_file", "input_src", "output_src"):
        getattr(L, n).argtypes = [vp, rp, cs_]
    L.input_data.argtypes = [vp, vp, rp, cs_]
    L.output_data.argtypes = [vp, vp, rp, cs_]
    L.argot_hip_set_max_sources.argtypes = [C.c_uint64]
    L.argot_hip_set_tile.argtypes = [C.c_int, C.c_int, C.c_int]
    L.argot_hip_set_graph.argtypes = [C.c_int]
    L.argot_hip_graph_captures.restype = C.c_int64
    L.argot_hip_autotune_step.argtypes = [C.c_void_p, C.c_void_p, rp, cm, cp, P(abi.TuneReport)]
    _lib = L
    return L


def check(what):
    """Raise if the last entry point reported an error."""
    e = lib().argot_hip_last_error()
    if e:
        msg = lib().argot_hip_error_string(e).decode()
        detail = lib().argot_hip_last_error_detail().decode()
        raise ArgotError(f"{what} failed: {msg} (code {e}) {detail}".rstrip())
