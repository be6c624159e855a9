"""ART diffuse transfer on the GPU against the oracle: the whole pass (rates
of every cell, crossing count) for several direction sets and mesh sizes up
to 64^3 with 768 directions, batch-size independence (also on the full
128^3 mesh), and the Kahan sums over directions."""
import numpy as np
import pytest

import oracle as O
from argot_amd import art as A
from argot_amd import ic

pytestmark = pytest.mark.gpu


def oracle_rates(w, nside, heatcool=True):
    g = O.art_grid(w.n, w.lunit, w.denstonh, w.helium_fact)
    return O.art_run(g, nside, w.dens, A.chem_array(w), A.UENETOK, heatcool)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    floor = 1e-6 * np.abs(b).max(0)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


@pytest.mark.parametrize("n,nside,heatcool,exp64", [(16, 2, True, True), (16, 8, True, True),
                                                    (16, 8, True, False), (32, 4, False, False)])
def test_diffuse_pass_matches_oracle(require_gpu, n, nside, heatcool, exp64):
    w = ic.art128(seed=11, n=n)
    print(f"art n={n} nside={nside} heatcool={heatcool} exp64={exp64}", flush=True)
    st = A.DiffuseStep(w, nside=nside, heatcool=heatcool, batch=16, exp64=exp64)
    stats = st.step()
    got = st.rates()
    ref, ncross = oracle_rates(w, nside, heatcool)
    assert stats["ndir"] == 12 * nside * nside
    assert stats["ncross"] == ncross
    assert np.isfinite(got).all()
    err = rel(got, ref)
    assert err.max() < 1e-5, err.max(0)
    st.close()


def test_batch_size_does_not_change_results(require_gpu):
    w = ic.art128(seed=12, n=16)
    a = A.DiffuseStep(w, nside=4, batch=1)
    a.step()
    ra = a.rates()
    a.close()
    b = A.DiffuseStep(w, nside=4, batch=64)
    b.step()
    np.testing.assert_array_equal(ra, b.rates())
    b.close()


def test_kahan_sums_over_directions(require_gpu):
    w = ic.art128(seed=13, n=16)
    p = A.DiffuseStep(w, nside=8, batch=32)
    p.step()
    rp = p.rates()
    p.close()
    k = A.DiffuseStep(w, nside=8, batch=32, kahan=True)
    k.step()
    rk = k.rates()
    k.close()
    ref, _ = oracle_rates(w, 8)
    assert rel(rk, ref).max() < 1e-5
    assert not np.array_equal(rk, rp)


def test_default_batch_matches_oracle_at_64(require_gpu):
    # the bench's settings (768 directions, automatic batch) on a 64^3 mesh
    w = ic.art128(seed=14, n=64)
    st = A.DiffuseStep(w, nside=8)
    stats = st.step()
    got = st.rates()
    st.close()
    ref, ncross = oracle_rates(w, 8)
    assert stats["ndir"] == 768 and stats["ncross"] == ncross
    assert np.isfinite(got).all()
    assert rel(got, ref).max() < 1e-5


def test_full_size_batches_agree(require_gpu):
    # 128^3 x 768 directions: the automatic batch (128 directions) and
    # batches of 32 give the same rates bit for bit and the same crossings
    w = ic.art128()
    a = A.DiffuseStep(w, nside=8)
    sa = a.step()
    ra = a.rates()
    a.close()
    b = A.DiffuseStep(w, nside=8, batch=32)
    sb = b.step()
    rb = b.rates()
    b.close()
    assert sa["ncross"] == sb["ncross"] > 0
    assert np.isfinite(ra).all() and (ra >= 0).all()
    np.testing.assert_array_equal(ra, rb)
