"""Segment builder (calc_ray_segment, src/calc_ray_segment.c:15-213) on one
rank, on the CPU: the oracle's domain loop pinned by its closed form, and
the library's segment_start (the code the fused kernel runs, built for the
host) against the oracle for sources inside, on and outside the box,
including sources far enough out that the reference's NSEG_PER_RAY assert
(:212) refuses the ray."""
import ctypes as C

import numpy as np

from argot_amd import radiation as R
from argot_amd._lib import lib
from helpers import cell_centre, oracle_problem, small_workload

NSEG_PER_RAY = 5
N = 16


def setup():
    w = small_workload(n=N, nsrc=1, seed=2)
    P = oracle_problem(w)
    freq, cs = R.freq_and_csect()
    run = R.run_param_for(w, freq, cs)
    return P, run


def product_start(run, src, ix, iy, iz):
    st = np.zeros(3, np.float32)
    s = np.ascontiguousarray(src, np.float32)
    n = lib().argot_hip_segment_start(C.byref(run), s.ctypes.data, ix, iy, iz, st.ctypes.data)
    return n, st


def shell_sources(rng, n, lo, hi):
    """Points in [lo, hi]^3 with at least one coordinate outside [0, 1]."""
    out = []
    while len(out) < n:
        p = rng.uniform(lo, hi, 3).astype(np.float32)
        if np.any(p < 0.0) or np.any(p > 1.0):
            out.append(p)
    return np.array(out)


def face_sources(rng):
    """Sources exactly on each face plane and just past each face."""
    pts = []
    for axis in range(3):
        for v in (0.0, 1.0, np.nextafter(np.float32(1.0), np.float32(2.0)),
                  np.nextafter(np.float32(0.0), np.float32(-1.0)), -0.25, 1.25):
            p = rng.uniform(0.1, 0.9, 3).astype(np.float32)
            p[axis] = v
            pts.append(p)
    return np.array(pts, np.float32)


def test_product_segment_start_matches_oracle():
    """Start point of the kept (last) segment bit for bit, and the segment
    count (the product counts up to NSEG_PER_RAY)."""
    P, run = setup()
    rng = np.random.default_rng(5)
    srcs = np.concatenate([rng.uniform(0.0, 1.0, (200, 3)).astype(np.float32),
                           shell_sources(rng, 400, -2.5, 3.5), face_sources(rng)])
    nmulti = 0
    for src in srcs:
        ix, iy, iz = (int(v) for v in rng.integers(0, N, 3))
        s, e = P.ray_segment(src, ix, iy, iz, maxseg=8)
        n, st = product_start(run, src, ix, iy, iz)
        assert n == min(len(s), NSEG_PER_RAY), (src, n, len(s))
        nmulti += len(s) > 1
        if n < NSEG_PER_RAY:
            assert np.array_equal(st.view(np.uint32), s[-1].view(np.uint32)), (src, st, s[-1])
    assert nmulti > 100   # the domain loop ran for many of them


def test_far_source_is_refused():
    """A source 20 box lengths out enters far from the box on another axis
    and needs more domain steps than NSEG_PER_RAY allows: the reference
    asserts, the library reports the ray as refused."""
    P, run = setup()
    src = np.array([-20.0, -0.5, 0.5], np.float32)
    s, _ = P.ray_segment(src, 8, 8, 8, maxseg=32)
    assert len(s) >= NSEG_PER_RAY
    n, _ = product_start(run, src, 8, 8, 8)
    assert n == NSEG_PER_RAY
    # just inside the limit: 4 segments are accepted
    for x in np.linspace(-8.0, -2.0, 61, dtype=np.float32):
        src = np.array([x, -0.5, 0.5], np.float32)
        s, _ = P.ray_segment(src, 8, 8, 8, maxseg=32)
        n, st = product_start(run, src, 8, 8, 8)
        assert n == min(len(s), NSEG_PER_RAY)
        if len(s) < NSEG_PER_RAY:
            assert np.array_equal(st.view(np.uint32), s[-1].view(np.uint32))


def test_oracle_domain_loop_closed_form():
    """One domain per axis: the entry point's truncated box coordinates
    (i, j, k) say how many unit planes separate it from the box, so the ray
    has 1 + |i| + |j| + |k| segments when no two planes are crossed at the
    same point (random sources never tie); the segments chain, and every
    interior end point lies on a plane of integer coordinate, on the line
    from source to target."""
    P, _ = setup()
    rng = np.random.default_rng(9)
    checked = 0
    for src in shell_sources(rng, 400, -3.0, 4.0):
        ix, iy, iz = (int(v) for v in rng.integers(0, N, 3))
        s, e = P.ray_segment(src, ix, iy, iz, maxseg=32)
        tgt = np.array([cell_centre(ix, N), cell_centre(iy, N), cell_centre(iz, N)], np.float64)
        assert np.array_equal(e[-1], tgt.astype(np.float32))
        for q in range(len(s) - 1):
            assert np.array_equal(e[q], s[q + 1])
            v = e[q].astype(np.float64)
            assert np.min(np.abs(v - np.round(v))) < 2e-6 * max(1.0, np.abs(v).max())
        u = tgt - src
        for p in np.concatenate([s, e]).astype(np.float64):
            w = p - src
            assert np.linalg.norm(np.cross(u, w)) <= 1e-5 * np.linalg.norm(u) * (
                np.linalg.norm(w) + 1.0)
        # the domain index of the entry point along each axis is its
        # truncated coordinate; each unit plane between it and the box is
        # one more segment (an entry exactly on the upper face is one
        # zero-length step)
        i = np.trunc(s[0].astype(np.float64)).astype(int)
        assert len(s) == 1 + int(np.abs(i).sum()), (src, s, i)
        checked += len(s) > 1
    assert checked > 50
