"""The step's HIP graph and the autotuner: graph replays (with eager work and
synchronisations between them) equal plain stream launches bit for bit, a
graph is captured again only when a launch argument changes, every tile
shape gives the same per-cell rates, so does the 64-bit cell addressing of
large meshes, and the autotuner reports and keeps the fastest candidate."""
import numpy as np
import pytest

from argot_amd import ic
from argot_amd import radiation as R
from argot_amd._lib import check, lib
from helpers import oracle_problem, rel_err, small_workload

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_defaults():
    yield
    L = lib()
    L.argot_hip_set_graph(1)
    L.argot_hip_set_tile(4, 4, 4)
    L.argot_hip_set_blocks_per_cu(24)
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


def thin_mesh_with_thick_cells(n=16, nsrc=6, seed=34):
    """Mostly neutral-poor gas (in-cell optical depth below 1 in every bin)
    with six dense cells (above 25): most 4x4x4 tiles hold thin cells only,
    the six that hold a dense cell mix both kinds, and other tile shapes
    group the same cells differently."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    dens = np.ones(shape, np.float32)
    for c in [(1, 2, 3), (6, 9, 4), (10, 5, 13), (13, 14, 1), (4, 12, 10), (9, 0, 8)]:
        dens[c] = 60.0
    fHI = np.full(shape, 0.05, np.float32)
    fHeI = np.full(shape, 0.05, np.float32)
    fHII = (1 - fHI).astype(np.float32)
    fHeII = (1 - fHeI).astype(np.float32)
    pos = ((rng.integers(0, n, (nsrc, 3)) + 0.5) / n).astype(np.float32)
    w = ic.Workload(f"thick{n}", n, dens, fHI, fHII, fHeI, fHeII, np.zeros(shape, np.float32),
                    pos, ic.power_law_rates(rng, nsrc), np.full(nsrc, 1.0e5, np.float32))
    # in-cell optical depth at the HI and HeI edges (6.3e-18 and 7.4e-18 cm^2)
    # over a chord of one to sqrt(3) cell sizes
    cell = w.lunit / n
    nh = dens.astype(np.float64) * w.denstonh
    depth = nh * cell * (fHI * 6.3e-18 + w.helium_fact * fHeI * 7.4e-18)
    assert (depth[dens == 1.0] * np.sqrt(3.0)).max() < 1.0
    assert depth[dens == 60.0].min() > 25.0
    return w


def test_thin_rays_beside_thick_ones(require_gpu, restore_defaults):
    """A thin ray's rates are the same bits whether its wavefront runs the
    thin loop alone or the combined loop for a thick neighbour, and both
    kinds of cells match the oracle."""
    L = lib()
    w = thin_mesh_with_thick_cells()
    st = R.PointSourceStep(w)
    L.argot_hip_set_tile(4, 4, 4)
    st.step()
    r0 = st.rates()
    for tile in [(1, 8, 8), (2, 2, 16), (8, 8, 1)]:
        L.argot_hip_set_tile(*tile)
        check("argot_hip_set_tile")
        st.step()
        np.testing.assert_array_equal(st.rates(), r0, err_msg=f"tile {tile}")
    r32, _, _ = oracle_problem(w).step(f64=True)
    fin = np.isfinite(r32)
    assert np.array_equal(fin, np.isfinite(r0))
    err = rel_err(r0[fin], r32[fin], 1e-30)
    assert err.max() < 1e-5, f"max rel err {err.max():.3e}"


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

