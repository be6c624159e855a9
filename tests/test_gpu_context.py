"""Contexts: each init_gpu makes one, found from the caller's cuda_mem_space,
so two meshes (and their streams) live in one process.  Interleaved and
concurrent (two threads) steps give the same bits as each mesh alone, and a
released context is refused."""
import ctypes as C
import threading

import numpy as np
import pytest

from argot_amd import radiation as R
from argot_amd._lib import lib
from helpers import ragged_workload, small_workload

pytestmark = pytest.mark.gpu

ERR_STATE = 3


def alone(w, tau_mode=0):
    s = R.PointSourceStep(w, tau_mode=tau_mode)
    st = s.step()
    r = s.rates()
    s.close()
    return st, r


def test_two_meshes_interleaved(require_gpu):
    wa = small_workload(n=16, nsrc=6, seed=11)
    wb = ragged_workload((12, 7, 5), nsrc=4, seed=24)
    sta, ra = alone(wa)
    stb, rb = alone(wb, tau_mode=1)
    a = R.PointSourceStep(wa)
    b = R.PointSourceStep(wb, tau_mode=1)
    assert a.cmem.mesh_dev != b.cmem.mesh_dev
    assert a.cpar.strm[0] != b.cpar.strm[0]
    for _ in range(2):
        s1 = a.step()
        s2 = b.step()
        assert s1["nray"] == sta["nray"] and s2["nray"] == stb["nray"]
        assert np.array_equal(a.rates().view(np.uint32), ra.view(np.uint32))
        assert np.array_equal(b.rates().view(np.uint32), rb.view(np.uint32))
    a.close()
    # b is untouched by a's release
    b.step()
    assert np.array_equal(b.rates().view(np.uint32), rb.view(np.uint32))
    b.close()


def test_two_threads_two_streams(require_gpu):
    ws = [small_workload(n=24, nsrc=9, seed=5), small_workload(n=20, nsrc=7, seed=6)]
    want = [alone(w)[1] for w in ws]
    steps = [R.PointSourceStep(w) for w in ws]
    errors = []
    got = [None, None]

    def run(i):
        try:
            for _ in range(3):
                steps[i].step()
            got[i] = steps[i].rates()
        except Exception as e:   # reported below
            errors.append(e)

    th = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th)
    assert not errors, errors
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint32), w.view(np.uint32))
    for s in steps:
        s.close()


def test_released_context_is_refused(require_gpu):
    w = small_workload(n=8, nsrc=2, seed=3)
    s = R.PointSourceStep(w)
    s.step()
    cmem = type(s.cmem)()
    C.memmove(C.addressof(cmem), C.addressof(s.cmem), C.sizeof(cmem))
    assert lib().argot_hip_release_gpu(C.byref(s.cmem)) == 0
    assert not s.cmem.mesh_dev
    # the old handle no longer names a context
    L = lib()
    L.step_radiation_tree(s.mesh, s.src, C.byref(s.run), None, C.byref(cmem), C.byref(s.cpar),
                          0.0)
    assert L.argot_hip_last_error() == ERR_STATE
    assert L.argot_hip_release_gpu(C.byref(cmem)) == ERR_STATE
