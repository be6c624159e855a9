"""ORACLE (test infrastructure) — ARGOT's on-disk formats restated with numpy,
to check the library's readers and writers byte for byte.

  mesh file "<prefix>_XXX_YYY_ZZZ" (output_mesh_single, src/output_data.c:14-45):
      struct run_param (5480 B here, as it lies in memory), then per local
      cell one struct fluid_mesh_io (src/fluid.h:118-125):
      dens eneg momx momy momz uene pot (float32) + struct prim_chem (13 float32)
  source file "<prefix>_src.dat" (output_src, src/output_data.c:60-73):
      uint64 nsrc, struct freq_param (3 + 256 float64), nsrc struct
      radiation_src (int32 type, float32 param, xpos, ypos, zpos, 4 B padding,
      256 float64 photon_rate)
"""
import numpy as np

PRIM_CHEM = np.dtype([(n, "<f4") for n in (
    "wmol", "felec", "fHI", "fHII", "fHeI", "fHeII", "fHeIII",
    "GammaHI", "HeatHI", "GammaHeI", "GammaHeII", "HeatHeI", "HeatHeII")])
MESH_IO = np.dtype([("dens", "<f4"), ("eneg", "<f4"), ("momx", "<f4"), ("momy", "<f4"),
                    ("momz", "<f4"), ("uene", "<f4"), ("pot", "<f4"), ("chem", PRIM_CHEM)])
FREQ_PARAM = np.dtype([("log_nu_min", "<f8"), ("log_nu_max", "<f8"), ("dlog_nu", "<f8"),
                       ("nu", "<f8", (256,))])
RADIATION_SRC = np.dtype({"names": ["type", "param", "xpos", "ypos", "zpos", "photon_rate"],
                          "formats": ["<i4", "<f4", "<f4", "<f4", "<f4", ("<f8", (256,))],
                          "offsets": [0, 4, 8, 12, 16, 24], "itemsize": 24 + 8 * 256})
RUN_PARAM_SIZE = 5480

assert MESH_IO.itemsize == 80 and FREQ_PARAM.itemsize == 2072
assert RADIATION_SRC.itemsize == 2072


def mesh_file_name(prefix, rank=(0, 0, 0)):
    return "%s_%03d_%03d_%03d" % (prefix, rank[0], rank[1], rank[2])


def write_mesh_file(path, run_bytes, records):
    assert len(run_bytes) == RUN_PARAM_SIZE and records.dtype == MESH_IO
    with open(path, "wb") as f:
        f.write(bytes(run_bytes))
        f.write(records.tobytes())


def read_mesh_file(path, ncell):
    with open(path, "rb") as f:
        head = f.read(RUN_PARAM_SIZE)
        body = f.read()
    assert len(head) == RUN_PARAM_SIZE and len(body) == ncell * MESH_IO.itemsize
    return head, np.frombuffer(body, dtype=MESH_IO)


def write_src_file(path, freq, src):
    assert freq.dtype == FREQ_PARAM and src.dtype == RADIATION_SRC
    with open(path, "wb") as f:
        f.write(np.uint64(len(src)).tobytes())
        f.write(freq.tobytes())
        f.write(src.tobytes())


def read_src_file(path):
    with open(path, "rb") as f:
        nsrc = int(np.frombuffer(f.read(8), "<u8")[0])
        freq = np.frombuffer(f.read(FREQ_PARAM.itemsize), FREQ_PARAM)[0]
        src = np.frombuffer(f.read(), RADIATION_SRC)
    assert len(src) == nsrc
    return nsrc, freq, src
