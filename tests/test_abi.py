"""The C-ABI library (CPU, no GPU calls): it loads, exports every function
include/argot_hip.h declares, and its structs have the reference's layout."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from argot_amd import abi
from argot_amd._lib import EXPORTS, LIB_PATH, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "argot_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[a-z][\w \*]*?\b(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


def test_library_exports_every_declared_symbol():
    L = lib()
    decl = declared_functions()
    assert len(decl) >= 20
    for name in decl + EXPORTS:
        assert hasattr(L, name), name


def test_struct_layout_matches_library():
    n = 32
    out = (C.c_int64 * n)()
    m = lib().argot_hip_abi_layout(out, n)
    assert list(out[:m]) == abi.layout_expected()


def test_reference_layout_sizes():
    """Sizes and offsets of the reference structs (gcc x86-64, __HELIUM__,
    NGRID_NU = 256), as the reference's own code sees them."""
    assert C.sizeof(abi.FluidMesh) == 156
    assert abi.FluidMesh.prev_chem.offset == 96
    assert C.sizeof(abi.RunParam) == 5480
    assert abi.RunParam.delta_x.offset == 5456
    assert C.sizeof(abi.RaySegment) == 56
    assert C.sizeof(abi.LightRay) == 2384
    assert C.sizeof(abi.LightRayIO) == 2096


def test_gpu_entry_points_fail_cleanly_without_init():
    """Entry points report ARGOT_ERR_STATE instead of aborting when the
    device context does not exist (no HIP call is made)."""
    L = lib()
    run = abi.RunParam()
    cm = abi.CudaMemSpace()
    cp = abi.CudaParam()
    L.zero_out_photoion_rate_cuda(C.byref(cm), C.byref(cp), C.byref(run))
    assert L.argot_hip_last_error() == 3
    assert L.argot_hip_error_string(3).decode().startswith("init_gpu")


@pytest.mark.parametrize("shape", [(0, 8, 8), (8, -1, 8), (4096, 4096, 1), (2048, 2048, 2048)])
def test_init_gpu_rejects_meshes_outside_the_index_range(shape):
    """Sizes are checked before any device call: empty meshes, nx*ny of 2^24
    or more (24-bit cell-index arithmetic) and more than 2^31-1 cells."""
    L = lib()
    run = abi.RunParam()
    run.nmesh_x_local, run.nmesh_y_local, run.nmesh_z_local = shape
    cm = abi.CudaMemSpace()
    cp = abi.CudaParam()
    L.init_gpu(None, C.byref(cm), C.byref(cp), C.byref(run))
    assert L.argot_hip_last_error() == 2    # ARGOT_ERR_ARG


def test_host_tree_handles_tiny_inputs():
    L = lib()
    box = np.array([0, 0, 0, 1, 1, 1], np.float32)
    pos = np.array([0.5, 0.5, 0.5], np.float32)
    spec = np.ones(256)
    h = L.argot_hip_tree_build(pos.ctypes.data, spec.ctypes.data, 1, 256, box.ctypes.data, 0.7)
    assert h
    t = np.array([0.1, 0.2, 0.3], np.float32)
    assert L.argot_hip_tree_rays_to(h, t.ctypes.data, 0, None, None, None) == 1
    L.argot_hip_tree_free(h)


def test_flat_dtype_views_match_ctypes():
    # the numpy view of a ctypes array reads and writes the same bytes,
    # tail padding included (RaySegment: 52 B of fields in a 56-B record)
    import ctypes as C
    a = (abi.RaySegment * 4)()
    v = abi.view_at(abi.RaySegment, C.addressof(a), 4)
    assert v.dtype.itemsize == C.sizeof(abi.RaySegment) == 56
    v["ray_indx"] = [1, 2, 3, 2**63 + 5]
    v["zpos_end"] = [0.5, 1.5, 2.5, 3.5]
    v["optical_depth_HeII"] = [7.0, 8.0, 9.0, 10.0]
    assert [r.ray_indx for r in a] == [1, 2, 3, 2**63 + 5]
    assert [r.zpos_end for r in a] == [0.5, 1.5, 2.5, 3.5]
    a[2].optical_depth_HI = 4.25
    assert v["optical_depth_HI"][2] == np.float32(4.25)
    assert list(v["optical_depth_HeII"]) == [7.0, 8.0, 9.0, 10.0]


def test_release_without_context_reports_state_and_keeps_last_error():
    """argot_hip_release_gpu returns its status and leaves the thread's last
    error alone (it may run from a destructor between a call and its
    check); statistics without a context are zero (no GPU call)."""
    L = lib()
    run = abi.RunParam()
    cm = abi.CudaMemSpace()
    cp = abi.CudaParam()
    L.zero_out_photoion_rate_cuda(C.byref(cm), C.byref(cp), C.byref(run))
    assert L.argot_hip_last_error() == 3
    assert L.argot_hip_release_gpu(C.byref(cm)) == 3
    assert L.argot_hip_release_gpu(None) == 2
    assert L.argot_hip_last_error() == 3
    st = abi.StepStats()
    L.argot_hip_step_stats(C.byref(st))
    assert st.nray == 0 and st.ms_kernel == 0.0
