"""Sources outside the box on the GPU (calc_ray_segment.c:107-192 with the
one-rank assign rule of accum_optical_depth.c:93-101): every ray's segment
start, optical depths and cell sequence against the oracle, the rates of
every cell, the NSEG_PER_RAY refusal (calc_ray_segment.c:212) reported as
ARGOT_ERR_TOO_MANY_SEGMENTS, and the one-rank guard of the fused step."""
import ctypes as C

import numpy as np
import pytest

from argot_amd import radiation as R
from argot_amd._lib import ArgotError, lib
from helpers import oracle_problem, rel_err, small_workload
from test_gpu_step import RATE_FLOOR, RATE_TOL, compare_records

pytestmark = pytest.mark.gpu

ERR_ARG = 2
ERR_TOO_MANY_SEGMENTS = 5

# outside each of the six faces, on a face, just past a face, and far
# enough out on two axes that the entry point needs two or three domain
# steps (three or four segments, within NSEG_PER_RAY = 5)
OUTSIDE = np.array([
    [-0.30, 0.52, 0.47], [1.40, 0.31, 0.62], [0.55, -0.21, 0.44],
    [0.23, 1.70, 0.81], [0.61, 0.42, -0.55], [0.33, 0.71, 1.30],
    [1.00, 0.40, 0.60], [np.nextafter(np.float32(1.0), np.float32(2.0)), 0.70, 0.20],
    [-2.60, -0.35, 0.52], [3.40, 0.45, -0.15], [0.50, -3.20, 1.60],
], np.float32)


def outside_workload(pos, n=16, seed=31):
    w = small_workload(n=n, nsrc=len(pos), seed=seed)
    w.src_pos = np.ascontiguousarray(pos, np.float32)
    return w


@pytest.mark.parametrize("n", [16, 12])
def test_sources_outside_every_face(require_gpu, n):
    # 16^3: power-of-two cells (the walk's face positions); 12^3: face counters
    w = outside_workload(OUTSIDE, n=n)
    P = oracle_problem(w)
    nseg = [len(P.ray_segment(p, 8, 8, 8)[0]) for p in OUTSIDE]
    assert max(nseg) >= 3   # the domain loop runs several steps
    step = R.PointSourceStep(w)
    rec, st = step.capture(400000)
    assert st["noverflow"] == 0
    assert st["nray"] == len(rec)
    shape = w.dens.shape
    compare_records(rec, P, shape)
    # the start of the kept segment, bit for bit
    nx, ny, nz = shape
    multi = 0
    for i in range(len(rec)):
        t = int(rec["target"][i])
        s, _ = P.ray_segment(rec["src"][i], t // (ny * nz), (t // nz) % ny, t % nz)
        multi += len(s) > 1
        assert np.array_equal(s[-1].view(np.uint32), rec["start"][i].view(np.uint32)), \
            f"ray {i}: start {rec['start'][i]} oracle {s[-1]}"
    assert multi > 0
    orc = P.ray_records(0, w.n ** 3)
    assert np.array_equal(orc["hash"], rec["hash"])
    assert np.array_equal(orc["ncross"], rec["ncross"])
    gpu = step.rates()
    r32, r64, ost = P.step(f64=True)
    assert ost["nray"] == st["nray"]
    fin = np.isfinite(r32)
    assert fin.all()   # no source sits in a target cell
    assert np.array_equal(fin, np.isfinite(gpu))
    err = rel_err(gpu, r32, RATE_FLOOR)
    assert err.max() < RATE_TOL, f"max rel err {err.max():.3e}"
    e64 = rel_err(gpu, r64, RATE_FLOOR)
    ref64 = rel_err(r32, r64, RATE_FLOOR)
    print(f"outside sources: max rel err vs FP64 {e64.max():.2e} "
          f"(reference arithmetic {ref64.max():.2e})")
    assert e64.max() <= ref64.max() + RATE_TOL
    assert e64.max() < 1e-3


def test_far_source_reports_too_many_segments(require_gpu):
    pos = np.array([[0.4, 0.5, 0.6], [-20.0, -0.5, 0.5]], np.float32)
    w = outside_workload(pos, n=8)
    step = R.PointSourceStep(w)
    with pytest.raises(ArgotError) as e:
        step.step()
    assert f"code {ERR_TOO_MANY_SEGMENTS}" in str(e.value)
    st = step.stats()
    assert st["noverflow"] > 0
    assert st["noverflow"] <= st["nray"]
    # without the far source the same step succeeds
    ok = R.PointSourceStep(outside_workload(pos[:1], n=8))
    assert ok.step()["noverflow"] == 0


@pytest.mark.parametrize("field,value", [("nnode_x", 2), ("nnode_z", 4), ("xmax_local", 0.5)])
def test_decomposed_rank_is_refused(require_gpu, field, value):
    """rank 0 of a decomposed mesh would treat the other ranks' regions as
    outside the box; the fused step refuses it."""
    w = small_workload(n=8, nsrc=3, seed=4)
    step = R.PointSourceStep(w)
    setattr(step.run, field, value)
    L = lib()
    L.step_radiation_tree(step.mesh, step.src, C.byref(step.run), None, C.byref(step.cmem),
                          C.byref(step.cpar), 0.0)
    assert L.argot_hip_last_error() == ERR_ARG
