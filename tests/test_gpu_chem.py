"""The chemistry on the GPU (step_chemistry, calc_dtime_gpu,
update_chemistry_gpu, and step_radiation_tree's iteration loop) against the
chemistry oracle on the same mesh, after a real point-source pass supplied
the rates.

The kernel and the oracle run the same float/double arithmetic; they differ
only where the device's double exp/log/pow differ from the host libm in the
last bit, which can move a float-rounded coefficient by one unit or flip a
bisection step.  So most cells agree bit for bit and the rest within the
bisection tolerance (1e-5 of the energy) carried into the fractions."""
import numpy as np
import pytest

import oracle as O
from argot_amd import abi
from argot_amd import radiation as R
from helpers import small_workload

pytestmark = pytest.mark.gpu

MESH_DTYPE = abi.mesh_view(abi.mesh_array(1)).dtype
FRACS = ("felec", "fHI", "fHII", "fHeI", "fHeII", "fHeIII")


def params(st):
    r = st.run
    return O.ChemParams(r.lunit, r.munit, r.tunit, r.uenetok, r.denstonh, st.w.helium_fact,
                        r.anow, r.znow)


def stepped(n=16, seed=41, nsrc=5):
    """A mesh after one point-source pass, with the (infinite) rates of the
    source cells set to zero so that every cell's chemistry is finite."""
    st = R.PointSourceStep(small_workload(n=n, nsrc=nsrc, seed=seed))
    st.step()
    st.mesh_bytes()
    v = abi.mesh_view(st.mesh)
    pc = v["prev_chem"]
    for f in R.RATE_FIELDS:
        x = pc[f]
        x[~np.isfinite(x)] = 0.0
        pc[f] = x
    v["prev_chem"] = pc
    v["chem"] = pc          # the 'previous iteration' the changes are measured against
    st.send_mesh()
    return st


def as_mesh(b):
    return np.frombuffer(b, dtype=MESH_DTYPE)


def bits(a):
    """Bit pattern of a (strided) float32 field."""
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("dtime", [1.0e-4, 3.0e-3])
def test_chemistry_step_matches_oracle(require_gpu, dtime):
    st = stepped()
    ref = np.frombuffer(st.mesh, np.uint8).copy()
    dc, du = st.chemistry(np.float32(dtime))
    cs = st.chem_stats()
    got = as_mesh(st.mesh_bytes())
    rdc, rdu, rns, rni = O.chem_step_mesh(ref, params(st), np.float32(dtime))
    want = as_mesh(ref)
    ncell = len(got)
    exact = np.ones(ncell, bool)
    for f in FRACS + R.RATE_FIELDS + ("wmol",):
        a, b = got["chem"][f], want["chem"][f]
        exact &= bits(a) == bits(b)
        np.testing.assert_allclose(a, b, rtol=2e-4, atol=1e-12, err_msg=f)
    a, b = got["uene"], want["uene"]
    exact &= bits(a) == bits(b)
    np.testing.assert_allclose(a, b, rtol=2e-5, err_msg="uene")
    assert np.all(got["duene"] == 0.0)
    # prev_chem / prev_uene are inputs and stay untouched
    assert got["prev_chem"].tobytes() == want["prev_chem"].tobytes()
    frac = exact.mean()
    print(f"dtime {dtime}: {frac:.4f} of {ncell} cells bit-exact; "
          f"sub-steps {cs['nsub']} (oracle {rns}), bisections {cs['niter']} (oracle {rni})")
    assert frac >= 0.9
    assert cs["ncell"] == ncell
    assert abs(cs["nsub"] - rns) <= 1e-3 * rns + 2
    np.testing.assert_allclose([dc, du], [max(rdc, np.float32(1.17549435e-38)),
                                          max(rdu, np.float32(1.17549435e-38))], rtol=2e-4)


def test_safe_step_matches_oracle(require_gpu):
    st = stepped()
    ref = np.frombuffer(st.mesh, np.uint8).copy()
    dt, dtc = st.chem_timestep()
    want = O.chem_dtime_mesh(ref, params(st))
    assert abs(dt - float(want)) <= 1e-6 * float(want), (dt, want)
    assert dtc == pytest.approx(dt / st.run.tunit, rel=1e-15)


def test_update_chemistry_matches_oracle_bit_for_bit(require_gpu):
    st = stepped()
    st.chemistry(np.float32(1e-4))
    ref = st.mesh_bytes()
    st.update_chemistry()
    got = st.mesh_bytes()
    O.chem_update_mesh(ref)
    assert np.array_equal(got, ref)
    m = as_mesh(got)
    assert m["prev_chem"].tobytes() == m["chem"].tobytes()
    assert np.array_equal(m["prev_uene"], m["uene"])


def test_iteration_loop_equals_its_calls(require_gpu):
    """step_radiation_tree with iterations = 5 equals the reference's loop
    written out with the entry points: pass, smooth, chemistry, repeated
    while a change exceeds 1e-3, then update_chemistry."""
    dtime = np.float32(2e-3)
    # sources off the cell centres: every rate is finite
    w = small_workload(n=12, nsrc=4, seed=43, inside_cells=False)
    a = R.PointSourceStep(w)
    sa = a.step(dtime=dtime, iterations=5)
    got = a.mesh_bytes()

    b = R.PointSourceStep(w)
    it = 0
    while True:
        b.step()
        b.smooth()
        dc, du = b.chemistry(dtime)
        it += 1
        if not ((dc > 1e-3 or du > 1e-3) and it < 5):
            break
    b.update_chemistry()
    want = b.mesh_bytes()
    assert sa["iterations"] == it
    assert 1 <= it <= 5
    # smooth() also writes the smoothed rates into the host mesh, the loop
    # does not: compare what lives on the device
    assert np.array_equal(got, want)


@pytest.mark.parametrize("refill", [1, 8, 64])
def test_queue_kernel_equals_one_cell_per_thread(require_gpu, refill):
    """The persistent cell-queue kernel (free lanes refilled once `refill`
    of them are free) and the one-cell-per-thread kernel give identical
    meshes and counts."""
    from argot_amd._lib import lib
    L = lib()
    st = stepped(n=16, seed=44)
    ref = np.frombuffer(st.mesh, np.uint8).copy()
    try:
        L.argot_hip_set_chem_queue(0)
        d0 = st.chemistry(np.float32(3e-3))
        c0 = st.chem_stats()
        m0 = st.mesh_bytes()
        np.copyto(np.frombuffer(st.mesh, np.uint8), ref)
        st.send_mesh()
        L.argot_hip_set_chem_queue(refill)
        d1 = st.chemistry(np.float32(3e-3))
        c1 = st.chem_stats()
        m1 = st.mesh_bytes()
    finally:
        L.argot_hip_set_chem_queue(1)
    assert np.array_equal(m0, m1)
    assert d0 == d1
    assert (c0["nsub"], c0["niter"]) == (c1["nsub"], c1["niter"])
