"""argot_hip_calc_optical_depth_noalloc (segments staged in the library's
fixed device array) against calc_optical_depth_cuda through the C-ABI: two
calls in a row, the second on a smaller segment list, each bit-identical
to the allocating entry point on the same segments; a list above the
array's limit is refused."""
import ctypes as C

import numpy as np
import pytest

from argot_amd import abi
from argot_amd import radiation as R
from argot_amd._lib import check, lib
from helpers import small_workload

pytestmark = pytest.mark.gpu

MAX_SEG = 1 << 20   # ARGOT_NOALLOC_MAX_SEGMENTS


def fill_segments(segp, s, e):
    seg = (abi.RaySegment * len(s)).from_address(segp.value)
    for i in range(len(s)):
        seg[i].ray_indx = i
        seg[i].xpos_start, seg[i].ypos_start, seg[i].zpos_start = (float(v) for v in s[i])
        seg[i].xpos_end, seg[i].ypos_end, seg[i].zpos_end = (float(v) for v in e[i])
        seg[i].optical_depth_HI = seg[i].optical_depth_HeI = seg[i].optical_depth_HeII = -1.0
    return seg


def taus(seg, n):
    return np.array([[seg[i].optical_depth_HI, seg[i].optical_depth_HeI,
                      seg[i].optical_depth_HeII] for i in range(n)], np.float32)


def run(L, step, entry, s, e):
    segp = C.c_void_p()
    L.allocate_pinned_segment(C.byref(segp), len(s))
    check("allocate_pinned_segment")
    try:
        seg = fill_segments(segp, s, e)
        step.run.nseg = len(s)
        getattr(L, entry)(segp, C.byref(step.cmem), C.byref(step.cpar), C.byref(step.run))
        check(entry)
        return taus(seg, len(s))
    finally:
        L.deallocate_pinned_segment(segp)


@pytest.mark.parametrize("kahan", [0, 1])
def test_noalloc_walk_equals_calc_optical_depth_cuda(require_gpu, kahan):
    w = small_workload(n=48, nsrc=3, seed=21)
    step = R.PointSourceStep(w)
    L = lib()
    rng = np.random.default_rng(17 + kahan)
    L.argot_hip_set_tau_mode(kahan)
    try:
        for count in (6000, 1500):   # the second call on a smaller list
            s = rng.uniform(-0.1, 1.1, (count, 3)).astype(np.float32)
            e = rng.uniform(0.0, 1.0, (count, 3)).astype(np.float32)
            k = count // 4   # centre-to-centre segments (exact ties)
            s[:k] = (rng.integers(0, w.n, (k, 3)) + 0.5) / w.n
            e[:k] = (rng.integers(0, w.n, (k, 3)) + 0.5) / w.n
            print(f"noalloc kahan={kahan} nseg={count}", flush=True)
            got = run(L, step, "argot_hip_calc_optical_depth_noalloc", s, e)
            ref = run(L, step, "calc_optical_depth_cuda", s, e)
            assert (got >= 0).all()
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), count
    finally:
        L.argot_hip_set_tau_mode(0)


def test_noalloc_refuses_more_than_its_array(require_gpu):
    w = small_workload(n=16, nsrc=2, seed=3)
    step = R.PointSourceStep(w)
    L = lib()
    segp = C.c_void_p()
    L.allocate_pinned_segment(C.byref(segp), 4)
    check("allocate_pinned_segment")
    try:
        step.run.nseg = MAX_SEG + 1   # refused before any segment is read
        L.argot_hip_calc_optical_depth_noalloc(segp, C.byref(step.cmem), C.byref(step.cpar),
                                              C.byref(step.run))
        assert L.argot_hip_last_error() == 2   # ARGOT_ERR_ARG
    finally:
        L.deallocate_pinned_segment(segp)
