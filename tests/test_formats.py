"""ARGOT's on-disk formats (mesh and source files of input_data.c /
output_data.c) through the library's C-ABI, against the format restated
with numpy (oracle/formats.py): what the library writes the restatement
reads back byte for byte, and what the restatement writes the library reads
into the reference's fields (prev_chem := chem, prev_uene := uene, other
fields untouched).  Header checks and short files are refused with an error
and leave the caller's arrays alone.  Host code only: no GPU."""
import ctypes as C
import os

import numpy as np
import pytest

import formats as F
from argot_amd import abi, ic
from argot_amd import radiation as R
from argot_amd._lib import lib
from helpers import small_workload


def state(n=6, nsrc=3, seed=3):
    """Host mesh / sources / run_param of a small workload (no device)."""
    w = small_workload(n=n, nsrc=nsrc, seed=seed)
    freq, cs = R.freq_and_csect()
    run = R.run_param_for(w, freq, cs)
    run.step, run.nspecies, run.nchannel = 7, 5, 3
    run.rank_x, run.rank_y, run.rank_z = 1, 2, 3
    src = R.blackbody_sources(w.src_pos, w.src_rate, w.src_temp, freq)
    mesh = R.fill_mesh(w)
    v = abi.mesh_view(mesh)
    rng = np.random.default_rng(seed)
    for f in ("eneg", "momx", "momy", "momz", "pot", "duene", "etrp"):
        v[f] = rng.normal(size=len(v)).astype(np.float32)
    return w, mesh, src, run


def run_bytes(run):
    return bytes(C.string_at(C.addressof(run), C.sizeof(run)))


def err():
    return lib().argot_hip_last_error()


# 48^3 and 41^3 cells pass through the library's 65536-record I/O buffer in
# two pieces, the second partial
@pytest.mark.parametrize("n", [6, 48])
def test_library_writes_the_reference_layout(tmp_path, n):
    w, mesh, src, run = state(n=n)
    prefix = str(tmp_path / "snap")
    R.write_snapshot(prefix, mesh, src, run)
    name = F.mesh_file_name(prefix, (1, 2, 3))
    assert os.path.getsize(name) == F.RUN_PARAM_SIZE + w.n ** 3 * 80
    head, rec = F.read_mesh_file(name, w.n ** 3)
    assert head == run_bytes(run)
    v = abi.mesh_view(mesh)
    for f in ("dens", "eneg", "momx", "momy", "momz", "uene", "pot"):
        assert np.array_equal(rec[f], v[f]), f
    assert rec["chem"].tobytes() == v["chem"].tobytes()
    nsrc, freq, s = F.read_src_file(prefix + "_src.dat")
    assert nsrc == w.nsrc
    assert freq.tobytes() == bytes(C.string_at(C.addressof(run.freq), C.sizeof(run.freq)))
    assert s.tobytes() == bytes(C.string_at(C.addressof(src), C.sizeof(abi.RadiationSrc) * nsrc))


@pytest.mark.parametrize("n", [6, 48])
def test_library_reads_what_the_format_says(tmp_path, n):
    w, mesh, src, run = state(n=n, seed=4)
    prefix = str(tmp_path / "ic")
    rng = np.random.default_rng(5)
    ncell = w.n ** 3
    rec = np.zeros(ncell, F.MESH_IO)
    for f in ("dens", "eneg", "momx", "momy", "momz", "uene", "pot"):
        rec[f] = rng.uniform(0.1, 2.0, ncell).astype(np.float32)
    for f in F.PRIM_CHEM.names:
        rec["chem"][f] = rng.uniform(0.0, 1.0, ncell).astype(np.float32)
    hb = bytearray(run_bytes(run))
    F.write_mesh_file(F.mesh_file_name(prefix, (1, 2, 3)), hb, rec)
    freq = np.zeros((), F.FREQ_PARAM)
    freq["nu"] = np.linspace(1.0, 30.0, 256)
    freq["dlog_nu"] = 0.25
    srcs = np.zeros(5, F.RADIATION_SRC)
    srcs["type"] = [0, 0, 1, 0, 1]
    srcs["xpos"] = rng.uniform(size=5)
    srcs["photon_rate"] = rng.uniform(size=(5, 256))
    F.write_src_file(prefix + "_src.dat", freq, srcs)

    m2 = abi.mesh_array(ncell)
    v2 = abi.mesh_view(m2)
    v2["duene"] = 9.5                 # not in the file: must stay
    v2["under_shock"] = 3
    s2 = (abi.RadiationSrc * 8)()
    r2 = abi.RunParam()
    r2.nmesh_x_local = r2.nmesh_y_local = r2.nmesh_z_local = w.n
    r2.rank_x, r2.rank_y, r2.rank_z = 1, 2, 3
    lib().input_data(m2, s2, C.byref(r2), prefix.encode())
    assert err() == 0
    for f in ("dens", "eneg", "momx", "momy", "momz", "uene", "pot"):
        assert np.array_equal(v2[f], rec[f]), f
    assert v2["chem"].tobytes() == rec["chem"].tobytes()
    assert v2["prev_chem"].tobytes() == rec["chem"].tobytes()
    assert np.array_equal(v2["prev_uene"], rec["uene"])
    assert np.all(v2["duene"] == 9.5) and np.all(v2["under_shock"] == 3)
    # the header is taken whole, except nsrc and freq which come from the source file
    got = bytearray(run_bytes(r2))
    want = bytearray(hb)
    off_n, off_f = abi.RunParam.nsrc.offset, abi.RunParam.freq.offset
    want[off_n:off_n + 8] = np.uint64(5).tobytes()
    want[off_f:off_f + F.FREQ_PARAM.itemsize] = freq.tobytes()
    assert got == want
    assert r2.nsrc == 5
    got_src = bytes(C.string_at(C.addressof(s2), C.sizeof(abi.RadiationSrc) * 5))
    assert got_src == srcs.tobytes()


@pytest.mark.parametrize("n", [5, 41])
def test_round_trip_through_the_library(tmp_path, n):
    w, mesh, src, run = state(n=n, seed=6)
    prefix = str(tmp_path / "rt")
    R.write_snapshot(prefix, mesh, src, run)
    m2, s2, r2 = R.read_snapshot(prefix, w.n, max_sources=16, rank=(1, 2, 3))
    v, v2 = abi.mesh_view(mesh), abi.mesh_view(m2)
    for f in ("dens", "eneg", "momx", "momy", "momz", "uene", "pot"):
        assert np.array_equal(v[f], v2[f])
    assert v2["chem"].tobytes() == v["chem"].tobytes()
    assert run_bytes(r2) == run_bytes(run)


@pytest.mark.parametrize("field,value", [("nmesh_x_local", 7), ("nspecies", 2),
                                         ("nchannel", 1), ("nmesh_z_total", 99)])
def test_header_mismatch_is_refused(tmp_path, field, value):
    w, mesh, src, run = state(seed=7)
    setattr(run, field, value)
    if field == "nmesh_x_local":
        run.nmesh_x_local = w.n          # written with the real size ...
    prefix = str(tmp_path / "bad")
    R.write_snapshot(prefix, mesh, src, run)
    m2 = abi.mesh_array(w.n ** 3)
    r2 = abi.RunParam()
    r2.nmesh_x_local = r2.nmesh_y_local = r2.nmesh_z_local = w.n
    r2.nmesh_x_total = r2.nmesh_y_total = r2.nmesh_z_total = w.n
    r2.rank_x, r2.rank_y, r2.rank_z = 1, 2, 3
    if field == "nmesh_x_local":
        r2.nmesh_x_local = value         # ... and read into a different one
    before = run_bytes(r2)
    lib().input_mesh(m2, C.byref(r2), prefix.encode())
    assert err() == 2        # ARGOT_ERR_ARG
    assert run_bytes(r2) == before
    assert np.all(abi.mesh_view(m2)["dens"] == 0.0)


def test_short_and_missing_files(tmp_path):
    w, mesh, src, run = state(seed=8)
    prefix = str(tmp_path / "short")
    R.write_snapshot(prefix, mesh, src, run)
    name = F.mesh_file_name(prefix, (1, 2, 3))
    with open(name, "r+b") as f:
        f.truncate(os.path.getsize(name) - 1)
    m2 = abi.mesh_array(w.n ** 3)
    r2 = abi.RunParam()
    r2.nmesh_x_local = r2.nmesh_y_local = r2.nmesh_z_local = w.n
    r2.rank_x, r2.rank_y, r2.rank_z = 1, 2, 3
    lib().input_mesh(m2, C.byref(r2), prefix.encode())
    assert err() == 6        # ARGOT_ERR_IO
    assert np.all(abi.mesh_view(m2)["dens"] == 0.0)
    lib().input_src(src, C.byref(r2), str(tmp_path / "none").encode())
    assert err() == 6


def test_too_many_sources_is_refused(tmp_path):
    w, mesh, src, run = state(nsrc=5, seed=9)
    prefix = str(tmp_path / "many")
    R.write_snapshot(prefix, mesh, src, run)
    L = lib()
    s2 = (abi.RadiationSrc * 4)()
    r2 = abi.RunParam()
    L.argot_hip_set_max_sources(4)
    try:
        L.input_src(s2, C.byref(r2), prefix.encode())
        assert err() == 2 and r2.nsrc == 0
        L.argot_hip_set_max_sources(5)
        s3 = (abi.RadiationSrc * 5)()
        L.input_src(s3, C.byref(r2), prefix.encode())
        assert err() == 0 and r2.nsrc == 5
    finally:
        L.argot_hip_set_max_sources(1024)


def test_huge_source_count_in_a_short_file_is_refused_before_allocating(tmp_path):
    """A source file whose header claims far more records than it holds is
    refused by its size (ARGOT_ERR_IO) even when the capacity allows the
    count: nothing the size of the claimed count is allocated, and nothing
    crosses the C-ABI as an exception."""
    w, mesh, src, run = state(nsrc=3, seed=10)
    prefix = str(tmp_path / "huge")
    R.write_snapshot(prefix, mesh, src, run)
    name = prefix + "_src.dat"
    with open(name, "r+b") as f:
        f.write(np.uint64(1 << 60).tobytes())     # nsrc field: 2^60 records
    L = lib()
    s2 = (abi.RadiationSrc * 3)()
    r2 = abi.RunParam()
    L.argot_hip_set_max_sources(1 << 62)
    try:
        L.input_src(s2, C.byref(r2), prefix.encode())
        assert err() == 6 and r2.nsrc == 0        # ARGOT_ERR_IO
    finally:
        L.argot_hip_set_max_sources(1024)
