"""ORACLE — test infrastructure, not product code.

ctypes binding of oracle/liboracle.so, the plain-C CPU restatement of ARGOT's
point-source radiation step.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py load this module; argot_amd never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# The same restatement built with the reference's own Makefile flags
# (-O3 -ffast-math -funroll-loops, src/Makefile:69): timed beside the IEEE
# build as the CPU baseline; never the checker.
LIB_FAST_PATH = os.path.join(HERE, "liboracle_fast.so")
REF_DIR = os.path.join(HERE, "_ref")

_libs = {}

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class Grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("xmin", C.c_float), ("ymin", C.c_float), ("zmin", C.c_float),
        ("xmax", C.c_float), ("ymax", C.c_float), ("zmax", C.c_float),
        ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
        ("lunit", C.c_double), ("denstonh", C.c_double),
        ("helium_fact", C.c_double),
        ("theta_crit", C.c_float),
        ("ngrid", C.c_int),
        ("nu", C.POINTER(C.c_double)),
        ("cs", C.POINTER(C.c_float)),
    ]


class Fields(C.Structure):
    _fields_ = [("dens", C.POINTER(C.c_float)), ("fHI", C.POINTER(C.c_float)),
                ("fHeI", C.POINTER(C.c_float)), ("fHeII", C.POINTER(C.c_float))]


class Stats(C.Structure):
    _fields_ = [("nray", C.c_uint64), ("ncross", C.c_uint64), ("nseg", C.c_uint64)]


def lib(fast=False):
    """Load liboracle.so (built by `make -C oracle`), or with fast=True the
    build at the reference's compiler flags (liboracle_fast.so)."""
    if fast in _libs:
        return _libs[fast]
    path = LIB_FAST_PATH if fast else LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `make -C oracle`")
    L = C.CDLL(path)
    L.orc_freq_grid.argtypes = [C.c_int, C.c_double, C.c_double, f64p, C.POINTER(C.c_double)]
    L.orc_photon_rate_bb.argtypes = [C.c_int, C.c_double, C.c_double, C.c_float, C.c_double, f64p]
    for n in ("orc_csect_HI", "orc_csect_HeI", "orc_csect_HeII"):
        getattr(L, n).argtypes = [C.c_double]
        getattr(L, n).restype = C.c_double
    L.orc_cross_sections.argtypes = [C.c_int, f64p, f32p]
    L.orc_tree_build.argtypes = [C.POINTER(Grid), C.c_int, f32p, f64p]
    L.orc_tree_build.restype = C.c_void_p
    L.orc_tree_free.argtypes = [C.c_void_p]
    L.orc_tree_nnode.argtypes = [C.c_void_p]
    L.orc_tree_node_members.argtypes = [C.c_void_p, C.c_int, i32p]
    L.orc_tree_node_members.restype = C.c_int
    L.orc_tree_node.argtypes = [C.c_void_p, C.c_int, f32p, C.POINTER(C.c_float),
                                C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.POINTER(C.c_double))]
    L.orc_count_ray_to.argtypes = [C.c_void_p, C.POINTER(Grid), C.c_int, C.c_int, C.c_int]
    L.orc_count_ray_to.restype = C.c_uint64
    L.orc_rays_to.argtypes = [C.c_void_p, C.POINTER(Grid), C.c_int, C.c_int, C.c_int,
                              C.c_uint64, i32p, i32p, f32p]
    L.orc_rays_to.restype = C.c_uint64
    L.orc_ray_spec.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.orc_ray_spec.restype = C.POINTER(C.c_double)
    L.orc_ray_segment.argtypes = [C.POINTER(Grid), f32p, C.c_int, C.c_int, C.c_int,
                                  C.c_int, f32p, f32p]
    L.orc_walk_fp32.argtypes = [C.POINTER(Grid), C.POINTER(Fields), f32p, f32p, f32p,
                                C.c_int, C.c_void_p]
    L.orc_walk_fp32_kahan.argtypes = [C.POINTER(Grid), C.POINTER(Fields), f32p, f32p, f32p]
    L.orc_walk_exact.argtypes = [C.POINTER(Grid), C.POINTER(Fields), f32p, f32p, f64p]
    L.orc_delta_length.argtypes = [C.POINTER(Grid), C.c_float, C.c_float, C.c_float]
    L.orc_delta_length.restype = C.c_float
    L.orc_ray_rate.argtypes = [C.POINTER(Grid), C.POINTER(Fields), f32p, f64p, C.c_int,
                               C.c_int, C.c_int, f64p, C.c_int, f64p]
    L.orc_step.argtypes = [C.POINTER(Grid), C.POINTER(Fields), C.c_void_p, C.c_int64,
                           C.c_int64, f32p, C.c_void_p, C.c_int, C.c_int, C.POINTER(Stats)]
    L.orc_ray_records.argtypes = [C.POINTER(Grid), C.POINTER(Fields), C.c_void_p,
                                  C.c_int64, C.c_int64, C.c_uint64, i64p, i32p, i32p,
                                  f32p, i32p, u64p]
    L.orc_ray_records.restype = C.c_uint64
    L.orc_thin_kernel.argtypes = [C.c_int, f64p, f32p, f64p, f64p, f64p]
    i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
    L.orc_pix2ang_ring.argtypes = [C.c_long, C.c_long, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
    L.orc_art_angles.argtypes = [C.c_int, C.c_int, f32p, i16p, i16p, i32p]
    L.orc_art_recomb.argtypes = [C.c_double, f64p]
    L.orc_art_step_factors.argtypes = [C.c_int, f32p]
    L.orc_art_rmesh.argtypes = [C.c_long, f32p, f32p, C.c_double, C.c_double, C.c_double,
                                f32p, f32p]
    L.orc_art_direction.argtypes = [C.POINTER(Grid), f32p, C.c_int, f32p, f32p, f32p, f32p]
    L.orc_art_direction.restype = C.c_long
    L.orc_art_gh_tot.argtypes = [C.c_long, f32p, f32p, f32p, f32p, f32p, f32p]
    L.orc_art_run.argtypes = [C.POINTER(Grid), C.c_int, C.c_int, f32p, f32p, C.c_double, f32p]
    L.orc_art_run.restype = C.c_long
    _libs[fast] = L
    return L


def freq_grid(ngrid=256, log_nu_min=0.0, log_nu_max=1.5):
    nu = np.zeros(ngrid)
    d = C.c_double()
    lib().orc_freq_grid(ngrid, log_nu_min, log_nu_max, nu, C.byref(d))
    return nu, d.value


def photon_rate_bb(temperature=1.0e5, total=1.0, ngrid=256, log_nu_min=0.0, log_nu_max=1.5):
    out = np.zeros(ngrid)
    lib().orc_photon_rate_bb(ngrid, log_nu_min, log_nu_max, temperature, total, out)
    return out


def cross_sections(nu):
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    cs = np.zeros(3 * len(nu), dtype=np.float32)
    lib().orc_cross_sections(len(nu), nu, cs)
    return cs.reshape(-1, 3)


class Problem:
    """A mesh + sources in the oracle's terms (box [0,1)^3 by default)."""

    def __init__(self, dens, fHI, fHeI, fHeII, src_pos, src_spec, *, lunit, denstonh,
                 helium_fact, theta=0.7, nu=None, cs=None, box=(0.0, 1.0), fast=False):
        L = self._L = lib(fast)
        self.shape = dens.shape
        nx, ny, nz = dens.shape
        self.dens = np.ascontiguousarray(dens, dtype=np.float32)
        self.fHI = np.ascontiguousarray(fHI, dtype=np.float32)
        self.fHeI = np.ascontiguousarray(fHeI, dtype=np.float32)
        self.fHeII = np.ascontiguousarray(fHeII, dtype=np.float32)
        if nu is None:
            nu, _ = freq_grid()
        self.nu = np.ascontiguousarray(nu, dtype=np.float64)
        self.cs = np.ascontiguousarray(cross_sections(self.nu) if cs is None else cs,
                                       dtype=np.float32).reshape(-1)
        lo, hi = box
        g = Grid()
        g.nx, g.ny, g.nz = nx, ny, nz
        g.xmin = g.ymin = g.zmin = lo
        g.xmax = g.ymax = g.zmax = hi
        g.dx = np.float32((np.float32(hi) - np.float32(lo)) / np.float32(nx))
        g.dy = np.float32((np.float32(hi) - np.float32(lo)) / np.float32(ny))
        g.dz = np.float32((np.float32(hi) - np.float32(lo)) / np.float32(nz))
        g.lunit, g.denstonh, g.helium_fact = lunit, denstonh, helium_fact
        g.theta_crit = theta
        g.ngrid = len(self.nu)
        g.nu = self.nu.ctypes.data_as(C.POINTER(C.c_double))
        g.cs = self.cs.ctypes.data_as(C.POINTER(C.c_float))
        self.grid = g
        f = Fields()
        f.dens = self.dens.ctypes.data_as(C.POINTER(C.c_float))
        f.fHI = self.fHI.ctypes.data_as(C.POINTER(C.c_float))
        f.fHeI = self.fHeI.ctypes.data_as(C.POINTER(C.c_float))
        f.fHeII = self.fHeII.ctypes.data_as(C.POINTER(C.c_float))
        self.fields = f
        self.src_pos = np.ascontiguousarray(src_pos, dtype=np.float32).reshape(-1)
        self.src_spec = np.ascontiguousarray(src_spec, dtype=np.float64).reshape(-1)
        self.nsrc = len(self.src_pos) // 3
        self.tree = L.orc_tree_build(C.byref(g), self.nsrc, self.src_pos, self.src_spec)

    def __del__(self):
        try:
            if self.tree:
                self._L.orc_tree_free(self.tree)
        except Exception:
            pass

    @property
    def ncell(self):
        return int(np.prod(self.shape))

    def count_ray_to(self, ix, iy, iz):
        return int(self._L.orc_count_ray_to(self.tree, C.byref(self.grid), ix, iy, iz))

    def rays_to(self, ix, iy, iz):
        L = self._L
        n = self.count_ray_to(ix, iy, iz)
        kind = np.zeros(max(n, 1), np.int32)
        ids = np.zeros(max(n, 1), np.int32)
        pos = np.zeros(3 * max(n, 1), np.float32)
        L.orc_rays_to(self.tree, C.byref(self.grid), ix, iy, iz, n, kind, ids, pos)
        return kind[:n], ids[:n], pos[:3 * n].reshape(-1, 3)

    def ray_spec(self, kind, ident):
        p = self._L.orc_ray_spec(self.tree, int(kind), int(ident))
        return np.ctypeslib.as_array(p, shape=(self.grid.ngrid,)).copy()

    def node(self, k):
        cm = np.zeros(3, np.float32)
        length, n, nch = C.c_float(), C.c_int(), C.c_int()
        sp = C.POINTER(C.c_double)()
        self._L.orc_tree_node(self.tree, k, cm, C.byref(length), C.byref(n), C.byref(nch),
                            C.byref(sp))
        spec = np.ctypeslib.as_array(sp, shape=(self.grid.ngrid,)).copy()
        return cm, length.value, n.value, nch.value, spec

    def node_members(self, k):
        """Original indices of the sources below node k (sorted-key order)."""
        out = np.zeros(max(self.nsrc, 1), np.int32)
        n = self._L.orc_tree_node_members(self.tree, k, out)
        return out[:n].copy()

    def nnode(self):
        return int(self._L.orc_tree_nnode(self.tree))

    def ray_segment(self, src, ix, iy, iz, maxseg=8):
        s = np.zeros(3 * maxseg, np.float32)
        e = np.zeros(3 * maxseg, np.float32)
        n = self._L.orc_ray_segment(C.byref(self.grid), np.ascontiguousarray(src, np.float32),
                                  ix, iy, iz, maxseg, s, e)
        return s[:3 * n].reshape(-1, 3), e[:3 * n].reshape(-1, 3)

    def walk(self, start, end, cells=False):
        tau = np.zeros(3, np.float32)
        cap = sum(self.shape) + 8
        buf = np.zeros(cap, np.int32)
        n = self._L.orc_walk_fp32(C.byref(self.grid), C.byref(self.fields),
                                np.ascontiguousarray(start, np.float32),
                                np.ascontiguousarray(end, np.float32), tau, cap,
                                buf.ctypes.data_as(C.c_void_p))
        if cells:
            return tau, n, buf[:min(n, cap)].copy()
        return tau, n

    def walk_kahan(self, start, end):
        tau = np.zeros(3, np.float32)
        n = self._L.orc_walk_fp32_kahan(C.byref(self.grid), C.byref(self.fields),
                                      np.ascontiguousarray(start, np.float32),
                                      np.ascontiguousarray(end, np.float32), tau)
        return tau, n

    def walk_exact(self, start, end):
        tau = np.zeros(3, np.float64)
        n = self._L.orc_walk_exact(C.byref(self.grid), C.byref(self.fields),
                                 np.ascontiguousarray(start, np.float32),
                                 np.ascontiguousarray(end, np.float32), tau)
        return tau, n

    def delta_length(self, dx, dy, dz):
        return self._L.orc_delta_length(C.byref(self.grid), dx, dy, dz)

    def ray_rate(self, src, spec, ix, iy, iz, tau, mode=0):
        out = np.zeros(6)
        self._L.orc_ray_rate(C.byref(self.grid), C.byref(self.fields),
                           np.ascontiguousarray(src, np.float32),
                           np.ascontiguousarray(spec, np.float64), ix, iy, iz,
                           np.ascontiguousarray(tau, np.float64), mode, out)
        return out

    def step(self, t0=0, t1=None, f64=False, kahan=False, nthreads=0):
        """Rates for targets [t0, t1): (rates_ref float32 [N,6], rates_f64 or None, stats)."""
        if t1 is None:
            t1 = self.ncell
        r32 = np.zeros(6 * self.ncell, np.float32)
        r64 = np.zeros(6 * self.ncell, np.float64) if f64 else None
        st = Stats()
        self._L.orc_step(C.byref(self.grid), C.byref(self.fields), self.tree, t0, t1, r32,
                       r64.ctypes.data_as(C.c_void_p) if f64 else None, int(kahan),
                       int(nthreads), C.byref(st))
        return (r32.reshape(-1, 6), None if r64 is None else r64.reshape(-1, 6),
                {"nray": st.nray, "ncross": st.ncross, "nseg": st.nseg})

    def ray_records(self, t0, t1):
        """Per-ray (target, kind, id, tau[3], ncross, cell-sequence hash)."""
        L = self._L
        n = sum(self.count_ray_to(*cell_index(im, self.shape)) for im in range(t0, t1))
        m = max(n, 1)
        tgt = np.zeros(m, np.int64)
        kind = np.zeros(m, np.int32)
        ids = np.zeros(m, np.int32)
        tau = np.zeros(3 * m, np.float32)
        nc = np.zeros(m, np.int32)
        h = np.zeros(m, np.uint64)
        w = L.orc_ray_records(C.byref(self.grid), C.byref(self.fields), self.tree, t0, t1,
                              n, tgt, kind, ids, tau, nc, h)
        assert w == n
        return {"target": tgt[:n], "kind": kind[:n], "id": ids[:n],
                "tau": tau[:3 * n].reshape(-1, 3), "ncross": nc[:n], "hash": h[:n]}


def cell_index(im, shape):
    nx, ny, nz = shape
    ix = im // (ny * nz)
    iy = (im - ix * ny * nz) // nz
    iz = im - ix * ny * nz - iy * nz
    return int(ix), int(iy), int(iz)


def thin_kernel(nu, cs, spec, tau_th):
    out = np.zeros(4)
    lib().orc_thin_kernel(len(nu), np.ascontiguousarray(nu, np.float64),
                          np.ascontiguousarray(cs, np.float32).reshape(-1),
                          np.ascontiguousarray(spec, np.float64),
                          np.ascontiguousarray(tau_th, np.float64), out)
    return out


# ---------------------------------------------------------------------------
# ART diffuse-photon transfer (art.c)

def pix2ang_ring(nside, ipix):
    th, ph = C.c_double(), C.c_double()
    lib().orc_pix2ang_ring(nside, ipix, C.byref(th), C.byref(ph))
    return th.value, ph.value


def art_angles(nside, sorted_=True):
    """(ovr [n,3] FP32, base_id [n], corner_id [n], corner_num [8])."""
    n = 12 * nside * nside
    ovr = np.zeros(3 * n, np.float32)
    base = np.zeros(n, np.int16)
    corner = np.zeros(n, np.int16)
    cnum = np.zeros(8, np.int32)
    lib().orc_art_angles(nside, 1 if sorted_ else 0, ovr, base, corner, cnum)
    return ovr.reshape(n, 3), base, corner, cnum


def art_step_factors(heatcool=True):
    f = np.zeros(6, np.float32)
    lib().orc_art_step_factors(1 if heatcool else 0, f)
    return f


def art_grid(n, lunit, denstonh, helium_fact, lo=0.0, hi=1.0):
    h = float(np.float32(np.float32(hi - lo) / np.float32(n)))
    g = Grid()
    g.nx = g.ny = g.nz = n
    g.xmin = g.ymin = g.zmin = lo
    g.xmax = g.ymax = g.zmax = hi
    g.dx = g.dy = g.dz = h
    g.lunit, g.denstonh, g.helium_fact = lunit, denstonh, helium_fact
    return g


def art_rmesh(dens, chem, denstonh, uenetok, helium_fact):
    """chem [ncell, 7] = felec, fHI, fHII, fHeI, fHeII, fHeIII, prev_uene."""
    ncell = dens.size
    kap = np.zeros(3 * ncell, np.float32)
    src = np.zeros(3 * ncell, np.float32)
    lib().orc_art_rmesh(ncell, np.ascontiguousarray(dens, np.float32).reshape(-1),
                        np.ascontiguousarray(chem, np.float32).reshape(-1), denstonh, uenetok,
                        helium_fact, kap, src)
    return kap.reshape(ncell, 3), src.reshape(ncell, 3)


def art_direction(g, ovr, base, kap, src):
    ncell = g.nx * g.ny * g.nz
    length = np.zeros(ncell, np.float32)
    I = np.zeros(3 * ncell, np.float32)
    nc = lib().orc_art_direction(C.byref(g), np.ascontiguousarray(ovr, np.float32), int(base),
                                 np.ascontiguousarray(kap, np.float32).reshape(-1),
                                 np.ascontiguousarray(src, np.float32).reshape(-1), length, I)
    return length, I.reshape(ncell, 3), nc


def art_run(g, nside, dens, chem, uenetok, heatcool=True):
    """Rates [ncell, 6] in mesh order added to zeros, and the crossings."""
    ncell = g.nx * g.ny * g.nz
    rates = np.zeros(6 * ncell, np.float32)
    nc = lib().orc_art_run(C.byref(g), nside, 1 if heatcool else 0,
                           np.ascontiguousarray(dens, np.float32).reshape(-1),
                           np.ascontiguousarray(chem, np.float32).reshape(-1), uenetok, rates)
    return rates.reshape(ncell, 6), nc


def smooth_rates(rates, shape):
    """smooth_photoion_rate on one rank: rates [ncell, 6] in mesh order."""
    L = lib()
    L.orc_smooth_rates.argtypes = [C.c_int, C.c_int, C.c_int, f32p, f32p]
    r = np.ascontiguousarray(rates, np.float32).reshape(-1)
    out = np.zeros_like(r)
    L.orc_smooth_rates(shape[0], shape[1], shape[2], r, out)
    return out.reshape(-1, 6)


# ---------------------------------------------------------------- chemistry

class ChemParams(C.Structure):
    """orc_chem_params: run_param's units and cosmology + helium fraction."""
    _fields_ = [("lunit", C.c_double), ("munit", C.c_double), ("tunit", C.c_double),
                ("uenetok", C.c_double), ("denstonh", C.c_double),
                ("helium_fact", C.c_double), ("anow", C.c_float), ("znow", C.c_float)]


CHEM_FIELDS = ("wmol", "felec", "fHI", "fHII", "fHeI", "fHeII", "fHeIII",
               "GammaHI", "HeatHI", "GammaHeI", "GammaHeII", "HeatHeI", "HeatHeII")
COOLING_NAMES = ("creh2", "crehe2", "crexihe2", "crehe3", "cioh1", "ciohe1", "ciohe2",
                 "cexh1", "cexhe1", "cexhe2", "brems", "compt")
_chem_bound = False


def _chem_lib():
    global _chem_bound
    L = lib()
    if not _chem_bound:
        P = C.POINTER
        cp = P(ChemParams)
        L.orc_chem_k.argtypes = [C.c_double, f64p]
        L.orc_chem_cooling.argtypes = [C.c_double, C.c_float, f64p]
        L.orc_chem_heatcool.argtypes = [f32p, C.c_float, C.c_double, C.c_double, C.c_double]
        L.orc_chem_heatcool.restype = C.c_double
        L.orc_chem_advance_heatcool.argtypes = [C.c_float, P(C.c_float), f32p, cp, C.c_float]
        L.orc_chem_dtime.argtypes = [C.c_float, C.c_float, f32p, cp]
        L.orc_chem_dtime.restype = C.c_float
        L.orc_chem_advance.argtypes = [C.c_float, C.c_float, P(C.c_float), f32p, cp, C.c_float,
                                       P(C.c_int), P(C.c_int)]
        u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
        L.orc_chem_step_mesh.argtypes = [u8p, C.c_int64, cp, C.c_float, P(C.c_float),
                                         P(C.c_float), P(C.c_uint64), P(C.c_uint64)]
        L.orc_chem_dtime_mesh.argtypes = [u8p, C.c_int64, cp]
        L.orc_chem_dtime_mesh.restype = C.c_float
        L.orc_chem_update_mesh.argtypes = [u8p, C.c_int64]
        _chem_bound = True
    return L


def chem_k(t):
    k = np.zeros(6)
    _chem_lib().orc_chem_k(float(t), k)
    return k


def chem_cooling(t, zred=0.0):
    o = np.zeros(12)
    _chem_lib().orc_chem_cooling(float(t), zred, o)
    return o


def chem_heatcool(chem13, zred, nH, T, helium_fact):
    return _chem_lib().orc_chem_heatcool(np.ascontiguousarray(chem13, np.float32), zred, nH, T,
                                         helium_fact)


def chem_advance_heatcool(dens, uene, chem13, p, dtime):
    """(new uene, bisection count)."""
    u = C.c_float(uene)
    n = _chem_lib().orc_chem_advance_heatcool(dens, C.byref(u),
                                              np.ascontiguousarray(chem13, np.float32),
                                              C.byref(p), dtime)
    return np.float32(u.value), n


def chem_dtime(dens, mesh_uene, chem13, p):
    return np.float32(_chem_lib().orc_chem_dtime(dens, mesh_uene,
                                                 np.ascontiguousarray(chem13, np.float32),
                                                 C.byref(p)))


def chem_advance(dens, mesh_uene, uene, chem13, p, dtime):
    """One cell: (new chem13, new uene, sub-steps, bisections)."""
    c = np.array(chem13, np.float32)
    u = C.c_float(uene)
    ns, ni = C.c_int(0), C.c_int(0)
    _chem_lib().orc_chem_advance(dens, mesh_uene, C.byref(u), c, C.byref(p), dtime,
                                 C.byref(ns), C.byref(ni))
    return c, np.float32(u.value), ns.value, ni.value


def chem_step_mesh(mesh_bytes, p, dtime):
    """step_chemistry_gpu on a copy-free uint8 view of struct fluid_mesh[]
    (modified in place): (diff_chem, diff_uene, nsub, niter)."""
    m = mesh_bytes.reshape(-1)
    dc, du = C.c_float(0), C.c_float(0)
    ns, ni = C.c_uint64(0), C.c_uint64(0)
    _chem_lib().orc_chem_step_mesh(m, m.size // 156, C.byref(p), dtime, C.byref(dc),
                                   C.byref(du), C.byref(ns), C.byref(ni))
    return np.float32(dc.value), np.float32(du.value), ns.value, ni.value


def chem_dtime_mesh(mesh_bytes, p):
    m = mesh_bytes.reshape(-1)
    return np.float32(_chem_lib().orc_chem_dtime_mesh(m, m.size // 156, C.byref(p)))


def chem_update_mesh(mesh_bytes):
    m = mesh_bytes.reshape(-1)
    _chem_lib().orc_chem_update_mesh(m, m.size // 156)


REF_ART = os.path.join(REF_DIR, "libref_art.so")
_ref_art = None


def ref_art():
    """The reference's own diffuse-photon recombination fits
    (oracle/_ref/libref_art.so), or None when it is not built."""
    global _ref_art
    if _ref_art is not None:
        return _ref_art
    if not os.path.exists(REF_ART):
        return None
    L = C.CDLL(REF_ART)
    L.ref_art_recomb.argtypes = [C.c_double, f64p]
    _ref_art = L
    return L


def art_recomb(t):
    """The oracle's recombination fits at t [K]: k02_A, k02_B, k04_A,
    k04_B, k06_A, k06_B."""
    out = np.zeros(6)
    lib().orc_art_recomb(float(t), out)
    return out


REF_CHEM = os.path.join(REF_DIR, "libref_chem.so")
_ref_chem = None


def ref_chem():
    """The reference's own CPU chemistry (oracle/_ref/libref_chem.so), or None
    when it is not built (it needs the reference sources)."""
    global _ref_chem
    if _ref_chem is not None:
        return _ref_chem
    if not os.path.exists(REF_CHEM):
        return None
    L = C.CDLL(REF_CHEM)
    P = C.POINTER
    L.ref_chem_helium_fact.restype = C.c_double
    L.ref_chem_k.argtypes = [C.c_double, f64p]
    L.ref_chem_cooling.argtypes = [C.c_double, C.c_float, f64p]
    L.ref_chem_heatcool.argtypes = [f32p, C.c_float, C.c_double, C.c_double]
    L.ref_chem_heatcool.restype = C.c_double
    L.ref_chem_advance_heatcool.argtypes = [C.c_float, C.c_float, P(C.c_float), f32p, f64p,
                                            C.c_float, C.c_float, C.c_float]
    L.ref_chem_dtime.argtypes = [C.c_float, C.c_float, f32p, f64p, C.c_float, C.c_float]
    L.ref_chem_dtime.restype = C.c_double
    L.ref_chem_advance.argtypes = [C.c_float, C.c_float, P(C.c_float), f32p, f64p, C.c_float,
                                   C.c_float, C.c_float]
    _ref_chem = L
    return L
