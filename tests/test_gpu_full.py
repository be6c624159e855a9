"""The step at full size: lognormal128 (128^3 lognormal mesh, 1024
power-law sources, compensated optical-depth sums) and clumped256 (256^3,
4096 sources in 8 clumps).  Every cell of eight target ranges spread over
the mesh and of four more at source cells is compared with the oracle
(relative 1e-5, the same inf/NaN pattern); the whole mesh is checked for
what holds at any size: a rerun is bit-identical, finite rates are
non-negative, and non-finite ones (rays whose source sits at the target
centre) are confined to a few cells."""
import numpy as np
import pytest

from argot_amd import ic
from argot_amd import radiation as R
from helpers import oracle_problem, rel_err

pytestmark = pytest.mark.gpu

RATE_TOL = 1e-5      # relative, against the oracle's reference arithmetic
RATE_FLOOR = 1e-30   # rates below this (s^-1 / erg s^-1) count as zero


def sampled_ranges(w, nrange=8, size=2048, nnear=4):
    """nrange target ranges spread over the mesh, and nnear more starting
    at source cells (where the rates are largest and the rays shortest)."""
    ncell = w.n ** 3
    starts = list(np.linspace(0, ncell - size, nrange).astype(np.int64))
    ijk = np.minimum((np.asarray(w.src_pos, np.float64) * w.n).astype(np.int64), w.n - 1)
    for i, j, k in ijk[:: max(1, len(ijk) // nnear)][:nnear]:
        starts.append(min((i * w.n + j) * w.n + k - size // 2, ncell - size))
    return [(int(max(s, 0)), int(max(s, 0)) + size) for s in starts]


@pytest.mark.parametrize("name,kahan", [("lognormal128", True), ("clumped256", False)])
def test_full_size_step(require_gpu, name, kahan):
    w = ic.WORKLOADS[name]()
    print(f"workload {name}: {w.n}^3, {w.nsrc} sources, kahan={kahan}", flush=True)
    step = R.PointSourceStep(w, tau_mode=1 if kahan else 0)
    st = step.step()
    a = step.rates()
    step.step()
    b = step.rates()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert st["ntarget"] == w.n ** 3 and st["nray"] > 0
    fin = np.isfinite(a)
    assert (a[fin] >= 0.0).all()
    assert (~fin.all(axis=1)).sum() <= w.nsrc
    P = oracle_problem(w)
    for t0, t1 in sampled_ranges(w):
        r32, _, _ = P.step(t0, t1, kahan=kahan)
        ref, got = r32[t0:t1], a[t0:t1]
        f = np.isfinite(ref)
        assert np.array_equal(f, np.isfinite(got)), f"targets [{t0}, {t1})"
        err = rel_err(got[f], ref[f], RATE_FLOOR)
        assert err.max() < RATE_TOL, f"targets [{t0}, {t1}): max rel err {err.max():.3e}"
