"""The reference's per-call GPU entry points through the C-ABI:
calc_optical_depth_cuda on host segment lists and calc_photoion_rate_cuda on
light_ray lists, against the CPU oracle."""
import ctypes as C

import numpy as np
import pytest

from argot_amd import abi
from argot_amd import radiation as R
from argot_amd._lib import check, lib
from helpers import cell_centre, library_node_spectra, oracle_problem, rel_err, small_workload

pytestmark = pytest.mark.gpu


def random_segments(rng, n, count, inside=True):
    if inside:
        s = rng.uniform(0.0, 1.0, (count, 3)).astype(np.float32)
        e = rng.uniform(0.0, 1.0, (count, 3)).astype(np.float32)
    else:
        s = rng.uniform(-0.2, 1.2, (count, 3)).astype(np.float32)
        e = rng.uniform(-0.2, 1.2, (count, 3)).astype(np.float32)
    # a share of lattice-aligned segments: centre to centre (exact ties),
    # axis-parallel, and zero-length
    k = count // 4
    a = (rng.integers(0, n, (k, 3)) + 0.5) / n
    b = (rng.integers(0, n, (k, 3)) + 0.5) / n
    s[:k] = a
    e[:k] = b
    e[k:k + 8] = s[k:k + 8]
    e[k + 8:k + 40, 1:] = s[k + 8:k + 40, 1:]
    return s, e


@pytest.mark.parametrize("kahan", [0, 1])
def test_calc_optical_depth_cuda(require_gpu, kahan):
    w = small_workload(n=48, nsrc=3, seed=21)
    P = oracle_problem(w)
    step = R.PointSourceStep(w)
    L = lib()
    rng = np.random.default_rng(5)
    s, e = random_segments(rng, w.n, 4000)
    segp = C.c_void_p()
    L.allocate_pinned_segment(C.byref(segp), len(s))
    check("allocate_pinned_segment")
    try:
        seg = (abi.RaySegment * len(s)).from_address(segp.value)
        for i in range(len(s)):
            seg[i].ray_indx = i
            seg[i].xpos_start, seg[i].ypos_start, seg[i].zpos_start = (float(v) for v in s[i])
            seg[i].xpos_end, seg[i].ypos_end, seg[i].zpos_end = (float(v) for v in e[i])
        step.run.nseg = len(s)
        L.argot_hip_set_tau_mode(kahan)
        L.calc_optical_depth_cuda(segp, C.byref(step.cmem), C.byref(step.cpar), C.byref(step.run))
        check("calc_optical_depth_cuda")
        got = np.array([[seg[i].optical_depth_HI, seg[i].optical_depth_HeI,
                         seg[i].optical_depth_HeII] for i in range(len(s))], np.float32)
    finally:
        L.argot_hip_set_tau_mode(0)
        L.deallocate_pinned_segment(segp)
    for i in range(len(s)):
        tau = P.walk_kahan(s[i], e[i])[0] if kahan else P.walk(s[i], e[i])[0]
        assert np.array_equal(tau.view(np.uint32), got[i].view(np.uint32)), (i, s[i], e[i])


def test_calc_photoion_rate_cuda(require_gpu):
    w = small_workload(n=16, nsrc=5, seed=8)
    P = oracle_problem(w)
    step = R.PointSourceStep(w)
    L = lib()
    n = w.n
    # every ray of every target, with the oracle's optical depths
    rays = []
    for im in range(n ** 3):
        ix, iy, iz = im // (n * n), (im // n) % n, im % n
        kind, ids, pos = P.rays_to(ix, iy, iz)
        for k, j, p in zip(kind, ids, pos):
            s, e = P.ray_segment(p, ix, iy, iz)
            tau, _ = P.walk(s[-1], e[-1])
            rays.append((im, ix, iy, iz, P.ray_spec(k, j), p, tau))
    nray = len(rays)
    arr = (abi.LightRay * nray)()
    for i, (im, ix, iy, iz, spec, p, tau) in enumerate(rays):
        r = arr[i]
        r.src.type = 0
        r.src.param = 1.0e5
        r.src.xpos, r.src.ypos, r.src.zpos = (float(v) for v in p)
        np.ctypeslib.as_array(r.src.photon_rate)[:] = spec
        r.ix_target, r.iy_target, r.iz_target = ix, iy, iz
        r.optical_depth_HI, r.optical_depth_HeI, r.optical_depth_HeII = (float(v) for v in tau)
    iop = C.c_void_p()
    L.allocate_pinned_light_ray_IO(C.byref(iop), nray)
    check("allocate_pinned_light_ray_IO")
    try:
        L.zero_out_photoion_rate_cuda(C.byref(step.cmem), C.byref(step.cpar), C.byref(step.run))
        check("zero_out_photoion_rate_cuda")
        step.run.nray = nray
        L.calc_photoion_rate_cuda(C.byref(step.cmem), arr, iop, C.byref(step.cpar),
                                  C.byref(step.run))
        check("calc_photoion_rate_cuda")
    finally:
        L.deallocate_pinned_light_ray_IO(iop)
    gpu = step.rates()
    r32, _, _ = P.step()
    fin = np.isfinite(r32)
    assert np.array_equal(fin, np.isfinite(gpu))
    err = rel_err(gpu[fin], r32[fin], 1e-30)
    assert err.max() < 1e-5, err.max()


def light_rays_from_records(rec, w, spectra):
    """light_ray[] of captured fused-step rays (target order, then ray
    order), each with its node's or source's spectrum row."""
    n = len(rec)
    arr = (abi.LightRay * n)()
    side = w.n
    t = rec["target"]
    L = lib()
    srcs = R.blackbody_sources(w.src_pos, w.src_rate, w.src_temp)
    src_spec = np.array([np.ctypeslib.as_array(srcs[i].photon_rate) for i in range(w.nsrc)])
    for i in range(n):
        r = arr[i]
        r.src.type = 0
        r.src.param = 1.0e5
        r.src.xpos, r.src.ypos, r.src.zpos = (float(v) for v in rec["src"][i])
        row = spectra[rec["id"][i]] if rec["kind"][i] == 0 else src_spec[rec["id"][i]]
        np.ctypeslib.as_array(r.src.photon_rate)[:] = row
        r.ix_target = int(t[i] // (side * side))
        r.iy_target = int((t[i] // side) % side)
        r.iz_target = int(t[i] % side)
        r.optical_depth_HI, r.optical_depth_HeI, r.optical_depth_HeII = (
            float(v) for v in rec["tau"][i])
    return arr


def rate_call(step, arr, n):
    L = lib()
    iop = C.c_void_p()
    L.allocate_pinned_light_ray_IO(C.byref(iop), n)
    check("allocate_pinned_light_ray_IO")
    try:
        L.zero_out_photoion_rate_cuda(C.byref(step.cmem), C.byref(step.cpar), C.byref(step.run))
        check("zero_out_photoion_rate_cuda")
        step.run.nray = n
        L.calc_photoion_rate_cuda(C.byref(step.cmem), arr, iop, C.byref(step.cpar),
                                  C.byref(step.run))
        check("calc_photoion_rate_cuda")
    finally:
        L.deallocate_pinned_light_ray_IO(iop)
    return step.rates()


def test_rate_entry_point_equals_fused_step(require_gpu):
    """Every ray the fused step forms, handed to calc_photoion_rate_cuda
    with its node's or source's spectrum: the per-target sums in ray order
    give the fused kernel's per-cell rates bit for bit (cells whose rays all
    take the FP32 loop); shuffling whole targets, or giving one target's
    rays in two runs, leaves every cell's sum unchanged."""
    w = small_workload(n=16, nsrc=9, seed=13)
    step = R.PointSourceStep(w)
    rec, st = step.capture(400000)
    fused = step.rates()
    spectra = library_node_spectra(w)
    arr = light_rays_from_records(rec, w, spectra)
    got = rate_call(step, arr, len(rec))
    ps = step.stats()
    assert 0 < ps["nspec_call"] < len(rec) and ps["ms_rate"] > 0.0
    fin = np.isfinite(fused).all(1)
    assert fin.sum() >= w.n ** 3 - w.nsrc
    assert np.array_equal(np.isfinite(got).all(1), fin)
    assert np.array_equal(got[fin].view(np.uint32), fused[fin].view(np.uint32))
    # targets in another order: the library regroups nothing within a
    # target, so each cell's sum is the same
    t = rec["target"]
    starts = np.flatnonzero(np.r_[True, t[1:] != t[:-1]])
    ends = np.r_[starts[1:], len(t)]
    perm = np.random.default_rng(3).permutation(len(starts))
    idx = np.concatenate([np.arange(starts[k], ends[k]) for k in perm])
    arr2 = light_rays_from_records(rec[idx], w, spectra)
    got2 = rate_call(step, arr2, len(idx))
    assert np.array_equal(got2[fin].view(np.uint32), fused[fin].view(np.uint32))
    # the second half of the longest target's rays moved to the end: the
    # library regroups them after the first half (a stable sort by target)
    k = int(np.argmax(ends - starts))
    mid = (starts[k] + ends[k]) // 2
    idx3 = np.r_[np.arange(0, mid), np.arange(ends[k], len(t)), np.arange(mid, ends[k])]
    got3 = rate_call(step, light_rays_from_records(rec[idx3], w, spectra), len(idx3))
    assert np.array_equal(got3[fin].view(np.uint32), fused[fin].view(np.uint32))
    # the rays read and the spectra numbered serially (the path taken when
    # two different spectra share a hash): the same bits and spectrum count
    L = lib()
    L.argot_hip_set_host_threads(0)
    try:
        got4 = rate_call(step, light_rays_from_records(rec[idx], w, spectra), len(idx))
        assert step.stats()["nspec_call"] == ps["nspec_call"]
    finally:
        L.argot_hip_set_host_threads(16)
    assert np.array_equal(got4[fin].view(np.uint32), fused[fin].view(np.uint32))


def test_merge_photoion_data(require_gpu):
    w = small_workload(n=8, nsrc=2, seed=1)
    step = R.PointSourceStep(w)
    L = lib()
    L.zero_out_photoion_rate_cuda(C.byref(step.cmem), C.byref(step.cpar), C.byref(step.run))
    check("zero_out_photoion_rate_cuda")
    v = abi.mesh_view(step.mesh)
    rng = np.random.default_rng(0)
    add = rng.random((w.n ** 3, 6)).astype(np.float32)
    for j, f in enumerate(R.RATE_FIELDS):
        v["prev_chem"][f] = add[:, j]
    L.merge_photoion_data(step.mesh, C.byref(step.cmem), C.byref(step.cpar), C.byref(step.run))
    check("merge_photoion_data")
    L.merge_photoion_data(step.mesh, C.byref(step.cmem), C.byref(step.cpar), C.byref(step.run))
    check("merge_photoion_data")
    got = step.rates()
    assert np.array_equal(got, add + add)


def test_walk_entry_point_equals_fused_walk(require_gpu):
    # every segment the fused step emits, walked again through
    # calc_optical_depth_cuda (one lane per segment): the same optical-depth
    # bits, and the kernel time is reported
    w = small_workload(n=32, nsrc=16, seed=31)
    step = R.PointSourceStep(w)
    rec, st = step.capture(1500000)
    n = len(rec)
    assert n == st["nray"] > 0
    side = w.n
    t = rec["target"]
    idx = np.stack([t // (side * side), (t // side) % side, t % side], axis=1)
    L = lib()
    segp = C.c_void_p()
    L.allocate_pinned_segment(C.byref(segp), n)
    check("allocate_pinned_segment")
    try:
        seg = abi.view_at(abi.RaySegment, segp.value, n)
        for k, a in enumerate("xyz"):
            seg[f"{a}pos_start"] = rec["start"][:, k]
            seg[f"{a}pos_end"] = cell_centre(idx[:, k], side)
        step.run.nseg = n
        L.calc_optical_depth_cuda(segp, C.byref(step.cmem), C.byref(step.cpar), C.byref(step.run))
        check("calc_optical_depth_cuda")
        tau = np.stack([seg["optical_depth_HI"], seg["optical_depth_HeI"],
                        seg["optical_depth_HeII"]], axis=1).copy()
    finally:
        L.deallocate_pinned_segment(segp)
    assert np.array_equal(tau.view(np.uint32), rec["tau"].view(np.uint32))
    assert step.stats()["ms_walk"] > 0.0


def test_photoion_data_round_trip(require_gpu):
    """recv_photoion_data gives the mesh's six rates per cell; send_photoion_data
    puts a table back into the device mesh and touches nothing else."""
    w = small_workload(n=12, nsrc=4, seed=41)
    step = R.PointSourceStep(w)
    step.step()
    want = step.rates()
    L = lib()
    n = w.n ** 3
    gamma = (abi.PhotoionRate * n)()
    L.recv_photoion_data(gamma, C.byref(step.cmem), C.byref(step.cpar), C.byref(step.run))
    check("recv_photoion_data")
    got = np.ctypeslib.as_array(gamma).view(np.float32).reshape(n, 6)
    np.testing.assert_array_equal(got, want)
    before = abi.mesh_view(step.mesh).copy()
    new = (np.arange(6 * n, dtype=np.float32).reshape(n, 6) + 0.5)
    np.ctypeslib.as_array(gamma).view(np.float32).reshape(n, 6)[:] = new
    L.send_photoion_data(gamma, C.byref(step.cmem), C.byref(step.cpar), C.byref(step.run))
    check("send_photoion_data")
    np.testing.assert_array_equal(step.rates(), new)
    after = abi.mesh_view(step.mesh)
    for f in after.dtype.names:
        if f != "prev_chem":
            assert after[f].tobytes() == before[f].tobytes(), f
    for f in after["prev_chem"].dtype.names:
        if f not in R.RATE_FIELDS:
            assert np.array_equal(after["prev_chem"][f], before["prev_chem"][f]), f
