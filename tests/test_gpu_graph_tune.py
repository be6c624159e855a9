"""The step's HIP graph and the autotuner: graph replays (with eager work and
synchronisations between them) equal plain stream launches bit for bit, a
graph is captured again only when a launch argument changes, every tile
shape gives the same per-cell rates, so does the 64-bit cell addressing of
large meshes, and the autotuner reports and keeps the fastest candidate."""
import numpy as np
import pytest

from argot_amd import radiation as R
from argot_amd._lib import check, lib
from helpers import small_workload

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_defaults():
    yield
    L = lib()
    L.argot_hip_set_graph(1)
    L.argot_hip_set_tile(4, 4, 4)
    L.argot_hip_set_blocks_per_cu(16)
    L.argot_hip_set_wide_index(0)


def test_graph_replays_equal_plain_launches(require_gpu, restore_defaults):
    L = lib()
    w = small_workload(n=16, nsrc=6, seed=31)
    st = R.PointSourceStep(w)
    L.argot_hip_set_graph(0)
    st.step()
    eager = st.rates()
    L.argot_hip_set_graph(1)
    c0 = L.argot_hip_graph_captures()
    for k in range(4):
        st.step()
        got = st.rates()            # eager readback and synchronisation between replays
        np.testing.assert_array_equal(got, eager, err_msg=f"replay {k}")
    assert L.argot_hip_graph_captures() == c0 + 1
    # a changed launch argument (tile shape) captures a new graph
    L.argot_hip_set_tile(2, 4, 8)
    st.step()
    assert L.argot_hip_graph_captures() == c0 + 2
    np.testing.assert_array_equal(st.rates(), eager)


@pytest.mark.parametrize("tile", [(2, 4, 8), (1, 8, 8), (8, 8, 1), (2, 2, 16)])
def test_tile_shapes_give_identical_rates(require_gpu, restore_defaults, tile):
    L = lib()
    w = small_workload(n=16, nsrc=6, seed=32)
    st = R.PointSourceStep(w)
    L.argot_hip_set_tile(4, 4, 4)
    s0 = st.step()
    r0 = st.rates()
    L.argot_hip_set_tile(*tile)
    check("argot_hip_set_tile")
    s1 = st.step()
    np.testing.assert_array_equal(st.rates(), r0)
    assert (s0["nray"], s0["ncross"]) == (s1["nray"], s1["ncross"])


def test_bad_tile_is_rejected(require_gpu, restore_defaults):
    L = lib()
    L.argot_hip_set_tile(4, 4, 8)
    assert L.argot_hip_last_error() == 2
    L.argot_hip_set_tile(3, 4, 4)
    assert L.argot_hip_last_error() == 2


def test_autotuner_reports_and_keeps_the_fastest(require_gpu, restore_defaults):
    w = small_workload(n=16, nsrc=6, seed=33)
    st = R.PointSourceStep(w)
    st.step()
    r0 = st.rates()
    rows, best = st.autotune()
    assert len(rows) == 24
    ms = [r["ms_kernel"] for r in rows]
    assert best == int(np.argmin(ms)) and all(m > 0 for m in ms)
    assert len({r["tile"] for r in rows}) == 8
    st.step()
    np.testing.assert_array_equal(st.rates(), r0)


@pytest.mark.parametrize("n", [24, 32])
@pytest.mark.parametrize("kahan", [0, 1])
def test_wide_cell_addresses_give_identical_results(require_gpu, restore_defaults, kahan, n):
    """The 64-bit cell-field addressing and face counters that meshes of
    2^28 cells or more take give the same rates and crossings as the 32-bit
    offsets, in the fused step (through a new graph) and in
    calc_optical_depth_cuda; on the 32^3 mesh (cell size 2^-5) the default
    form also keeps running face positions instead of face counters."""
    import ctypes as C
    from argot_amd import abi
    L = lib()
    w = small_workload(n=n, nsrc=9, seed=33)
    st = R.PointSourceStep(w, tau_mode=kahan)
    s0 = st.step()
    r0 = st.rates()
    c0 = L.argot_hip_graph_captures()
    L.argot_hip_set_wide_index(1)
    s1 = st.step()
    assert L.argot_hip_graph_captures() == c0 + 1
    np.testing.assert_array_equal(st.rates(), r0)
    assert (s0["nray"], s0["ncross"]) == (s1["nray"], s1["ncross"])
    rng = np.random.default_rng(7)
    n = 3000
    segp = C.c_void_p()
    L.allocate_pinned_segment(C.byref(segp), n)
    check("allocate_pinned_segment")
    try:
        seg = abi.view_at(abi.RaySegment, segp.value, n)
        for k, a in enumerate("xyz"):
            seg[f"{a}pos_start"] = rng.uniform(0.0, 1.0, n).astype(np.float32)
            seg[f"{a}pos_end"] = rng.uniform(0.0, 1.0, n).astype(np.float32)
        st.run.nseg = n
        taus = []
        for wide in (1, 0):
            L.argot_hip_set_wide_index(wide)
            L.argot_hip_set_tau_mode(kahan)
            L.calc_optical_depth_cuda(segp, C.byref(st.cmem), C.byref(st.cpar), C.byref(st.run))
            check("calc_optical_depth_cuda")
            taus.append(np.stack([seg["optical_depth_HI"], seg["optical_depth_HeI"],
                                  seg["optical_depth_HeII"]], axis=1).copy())
    finally:
        L.argot_hip_set_tau_mode(0)
        L.deallocate_pinned_segment(segp)
    assert np.array_equal(taus[0].view(np.uint32), taus[1].view(np.uint32))

