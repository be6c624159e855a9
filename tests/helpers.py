"""Shared set-up for the parity tests: one workload, given to both the HIP
library (through its C-ABI) and the CPU oracle, with identical inputs (the
run_param tables and the radiation_src spectra are inputs of the step)."""
import numpy as np

import oracle as O
from argot_amd import ic
from argot_amd import radiation as R


def spectra_of(srcs, n):
    return np.array([np.ctypeslib.as_array(srcs[i].photon_rate) for i in range(n)])


def tables():
    freq, cs = R.freq_and_csect()
    nu = np.ctypeslib.as_array(freq.nu).copy()
    csa = np.array([[c.csect_HI, c.csect_HeI, c.csect_HeII] for c in cs], dtype=np.float32)
    return freq, cs, nu, csa


def oracle_problem(w, theta=0.7, fast=False):
    """The workload in the oracle's terms; fast=True: the oracle built at the
    reference's compiler flags (timing only, never a checker)."""
    freq, cs, nu, csa = tables()
    srcs = R.blackbody_sources(w.src_pos, w.src_rate, w.src_temp, freq)
    spec = spectra_of(srcs, w.nsrc)
    return O.Problem(w.dens, w.fHI, w.fHeI, w.fHeII, w.src_pos, spec, lunit=w.lunit,
                     denstonh=w.denstonh, helium_fact=w.helium_fact, theta=theta, nu=nu,
                     cs=csa, fast=fast)


def small_workload(n=16, nsrc=5, seed=7, lognormal=True, inside_cells=True):
    """A tiny mesh for exhaustive checks (every target, every ray)."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    if lognormal:
        dens = ic.lognormal_density(n, seed + 99)
        fHI = rng.uniform(1e-4, 1.0, shape).astype(np.float32)
        fHeI = rng.uniform(1e-4, 1.0, shape).astype(np.float32)
    else:
        dens = np.ones(shape, np.float32)
        fHI = np.full(shape, 0.999, np.float32)
        fHeI = np.full(shape, 0.999, np.float32)
    fHII = (1 - fHI).astype(np.float32)
    fHeII = (1 - fHeI).astype(np.float32)
    fHeIII = np.zeros(shape, np.float32)
    if inside_cells:
        pos = ((rng.integers(0, n, (nsrc, 3)) + 0.5) / n).astype(np.float32)
    else:
        pos = rng.uniform(0.0, 1.0, (nsrc, 3)).astype(np.float32)
    rates = ic.power_law_rates(rng, nsrc)
    return ic.Workload(f"small{n}", n, dens, fHI, fHII, fHeI, fHeII, fHeIII, pos, rates,
                       np.full(nsrc, 1.0e5, np.float32))


def ragged_workload(shape, nsrc=4, seed=17):
    """A non-cubic mesh (sizes that are not multiples of the 4x4x4 target
    tile, non-cubic cells) with sources off the cell centres."""
    rng = np.random.default_rng(seed)
    dens = np.exp(rng.normal(0.0, 0.7, shape)).astype(np.float32)
    fHI = rng.uniform(1e-4, 1.0, shape).astype(np.float32)
    fHeI = rng.uniform(1e-4, 1.0, shape).astype(np.float32)
    fHII = (1 - fHI).astype(np.float32)
    fHeII = (1 - fHeI).astype(np.float32)
    pos = rng.uniform(0.02, 0.98, (nsrc, 3)).astype(np.float32)
    rates = ic.power_law_rates(rng, nsrc) if nsrc else np.zeros(0)
    return ic.Workload("ragged", shape[0], dens, fHI, fHII, fHeI, fHeII,
                       np.zeros(shape, np.float32), pos, rates,
                       np.full(nsrc, 1.0e5, np.float32))


def rel_err(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def cell_centre(idx, n, lo=0.0, hi=1.0):
    """Cell centre as the reference forms it: lo + ((float)i + 0.5)*delta in
    double, rounded to float (setup_light_ray_tree.c:341-343)."""
    d = np.float64(np.float32(np.float32(hi - lo) / np.float32(n)))
    return np.float32(np.float64(np.float32(lo)) + (np.float64(np.float32(idx)) + 0.5) * d)


def library_node_spectra(w, theta=0.7):
    """Spectra [nnode, 256] of the library's own source tree (its node
    labels, which are the `id` of kind-0 ray records)."""
    import ctypes as C
    from argot_amd._lib import lib
    freq, _, _, _ = tables()
    srcs = R.blackbody_sources(w.src_pos, w.src_rate, w.src_temp, freq)
    spec = np.ascontiguousarray(spectra_of(srcs, w.nsrc))
    pos = np.ascontiguousarray(w.src_pos, np.float32)
    L = lib()
    box = (C.c_float * 6)(0, 0, 0, 1, 1, 1)
    h = L.argot_hip_tree_build(pos.ctypes.data, spec.ctypes.data, w.nsrc, spec.shape[1], box,
                               theta)
    try:
        nn = L.argot_hip_tree_nnode(h)
        return np.array([np.ctypeslib.as_array(L.argot_hip_tree_spec(h, u), shape=(spec.shape[1],))
                         for u in range(nn)])
    finally:
        L.argot_hip_tree_free(h)


RATE_NAMES = ("GammaHI", "HeatHI", "GammaHeI", "GammaHeII", "HeatHeI", "HeatHeII")


def fp64_rate_error(x, r64, floor=1e-30):
    """Per rate: (max, median) relative error of x [N, 6] against the FP64
    oracle's rates, over cells where both are finite."""
    x = np.asarray(x, np.float64)
    fin = np.isfinite(r64).all(1) & np.isfinite(x).all(1)
    e = rel_err(x[fin], r64[fin], floor)
    return {n: (float(e[:, k].max()), float(np.median(e[:, k]))) for k, n in enumerate(RATE_NAMES)}


def fp64_tau_error(rec, P, shape):
    """(max, median) relative error of the captured rays' FP32 optical
    depths (3 species) against the FP64 oracle's exact-chord walk over the
    same segment (rays with zero optical depth in both count as exact)."""
    nx, ny, nz = shape
    errs = np.zeros((len(rec), 3))
    for i in range(len(rec)):
        t = int(rec["target"][i])
        e = np.array([cell_centre(v, n) for v, n in
                      zip((t // (ny * nz), (t // nz) % ny, t % nz), shape)], np.float32)
        t64, _ = P.walk_exact(rec["start"][i], e)
        t32 = rec["tau"][i].astype(np.float64)
        d = np.abs(t32 - t64)
        errs[i] = np.where(d == 0.0, 0.0, d / np.maximum(np.abs(t64), 1e-30))
    return float(errs.max()), float(np.median(errs))


def report_fp64(label, gpu, r32, r64, tau=None):
    """Print the FP64 errors of the GPU rates beside those of the
    reference's own arithmetic (the oracle's FP32 path); returns both."""
    g = fp64_rate_error(gpu, r64)
    r = fp64_rate_error(r32, r64)
    for n in RATE_NAMES:
        print(f"{label} {n}: vs FP64 max {g[n][0]:.2e} median {g[n][1]:.2e} "
              f"(reference arithmetic max {r[n][0]:.2e} median {r[n][1]:.2e})")
    if tau is not None:
        print(f"{label} tau: vs FP64 max {tau[0]:.2e} median {tau[1]:.2e}")
    return g, r
