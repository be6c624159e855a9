"""Rate surrogate (host side): training, evaluation and inspection of the
3 -> 64 -> 64 -> 6 network behind argot_hip_set_rate_mode(1).

The network maps log10 of the threshold optical depths (HI, HeI, HeII) of a
ray to log10 of the unit-spectrum kernels F_s, G_s (s = HI, HeI, HeII); the
photon-conserving per-ray rates follow from their 2-point Gauss-Legendre mean
along the in-cell optical-depth path (calc_photoion_rate.c:285-387 in closed
form).  Everything numerical runs in libargot_hip.so on the GPU.
"""
import ctypes as C
import math

import numpy as np

from . import abi
from ._lib import check, lib

NOUT = 6
OUTPUTS = ("F_HI", "F_HeI", "F_HeII", "G_HI", "G_HeI", "G_HeII")

# 128 log bins from 13.6 eV to 1 keV (nu / nu_L from 1 to 1000/13.6).
C4_LOG_NU_MAX = math.log10(1000.0 / 13.6)


def spectral_grid(ngrid, log_nu_max, temperature=1.0e5, log_nu_min=0.0):
    """(nu [ngrid], cs [ngrid*3] FP32, spec [ngrid] unit total) of a log grid
    with a blackbody spectrum, from the library's table code."""
    nu = np.zeros(ngrid)
    cs = np.zeros(3 * ngrid, dtype=np.float32)
    spec = np.zeros(ngrid)
    lib().argot_hip_spectral_grid(ngrid, log_nu_min, log_nu_max, temperature,
                                  nu.ctypes.data, cs.ctypes.data, spec.ctypes.data)
    return nu, cs, spec


def step_grid(freq, csect, spec):
    """Tables of the step's own frequency grid: freq_param.nu as the FP32
    values the rate kernels use, run_param.csect, and one source spectrum."""
    ng = abi.NGRID_NU
    nu = np.array([float(np.float32(freq.nu[k])) for k in range(ng)])
    cs = np.array([[c.csect_HI, c.csect_HeI, c.csect_HeII] for c in csect[:ng]],
                  dtype=np.float32).reshape(-1)
    return nu, cs, np.asarray(spec, dtype=np.float64)[:ng].copy()


class Config:
    """Training configuration (argot_surrogate_config)."""

    def __init__(self, nu, cs, spec, nsample=1 << 22, batch=16384, nstep=20000,
                 lr_max=2e-3, lr_min=1e-6, seed=12345, log_tau_lo=-6.0, log_tau_hi=4.0,
                 trust_n=32, trust_tol=1e-3, dtau_max=0.8, seeds=None, jitter=0.1,
                 poly_degree=0, nvalidate=0):
        """seeds: log10 threshold optical depths [m,3] the samples are drawn
        around (Gaussian jitter of std `jitter`); None: uniform over the
        hull.  poly_degree 3: cubic least-squares baseline under the
        network."""
        self.nu = np.ascontiguousarray(nu, dtype=np.float64)
        self.cs = np.ascontiguousarray(cs, dtype=np.float32)
        self.spec = np.ascontiguousarray(spec, dtype=np.float64)
        self.seeds = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.float32)
        self.c = abi.SurrogateConfig(
            ngrid=len(self.nu),
            nu=self.nu.ctypes.data_as(C.POINTER(C.c_double)),
            cs=self.cs.ctypes.data_as(C.POINTER(C.c_float)),
            spec=self.spec.ctypes.data_as(C.POINTER(C.c_double)),
            nsample=nsample, batch=batch, nstep=nstep, lr_max=lr_max, lr_min=lr_min,
            seed=seed, log_tau_lo=log_tau_lo, log_tau_hi=log_tau_hi, trust_n=trust_n,
            trust_tol=trust_tol, dtau_max=dtau_max,
            seed_log_tau=(None if self.seeds is None
                          else self.seeds.ctypes.data_as(C.POINTER(C.c_float))),
            nseed=0 if self.seeds is None else len(self.seeds), jitter=jitter,
            poly_degree=poly_degree, nvalidate=nvalidate)


def train(cfg):
    """Generate samples, train on the GPU, validate; returns the report."""
    rep = abi.SurrogateReport()
    lib().argot_hip_surrogate_train(C.byref(cfg.c), C.byref(rep))
    check("argot_hip_surrogate_train")
    return {k: getattr(rep, k) for k, _ in abi.SurrogateReport._fields_}


def samples(cfg, n):
    """First n training samples: (log10 tau [n,3], log10 kernels [n,6])."""
    x = np.zeros((n, 3), dtype=np.float32)
    y = np.zeros((n, NOUT), dtype=np.float32)
    lib().argot_hip_surrogate_samples(C.byref(cfg.c), n, x.ctypes.data, y.ctypes.data)
    check("argot_hip_surrogate_samples")
    return x, y


def get_params():
    p = np.zeros(abi.SURROGATE_NPARAM, dtype=np.float32)
    norm = np.zeros(2 + 2 * NOUT, dtype=np.float32)
    lib().argot_hip_surrogate_get_params(p.ctypes.data, norm.ctypes.data)
    check("argot_hip_surrogate_get_params")
    return p, norm


def set_params(p, norm):
    p = np.ascontiguousarray(p, dtype=np.float32)
    norm = np.ascontiguousarray(norm, dtype=np.float32)
    lib().argot_hip_surrogate_set_params(p.ctypes.data, norm.ctypes.data)
    check("argot_hip_surrogate_set_params")


def evaluate(log_tau):
    """Network outputs (log10 kernels [n,6]) at log10 tau [n,3]."""
    x = np.ascontiguousarray(log_tau, dtype=np.float32)
    y = np.zeros((len(x), NOUT), dtype=np.float32)
    lib().argot_hip_surrogate_eval(x.ctypes.data, y.ctypes.data, len(x))
    check("argot_hip_surrogate_eval")
    return y


def gradient(log_tau, log_kernel):
    """(gradient [NPARAM], loss) of the normalised MSE, training kernels."""
    x = np.ascontiguousarray(log_tau, dtype=np.float32)
    y = np.ascontiguousarray(log_kernel, dtype=np.float32)
    g = np.zeros(abi.SURROGATE_NPARAM, dtype=np.float32)
    loss = C.c_double()
    lib().argot_hip_surrogate_gradient(x.ctypes.data, y.ctypes.data, len(x), g.ctypes.data,
                                       C.byref(loss))
    check("argot_hip_surrogate_gradient")
    return g, loss.value


def trust_error(trust_n):
    e = np.zeros(trust_n ** 3, dtype=np.float32)
    lib().argot_hip_surrogate_trust_error(e.ctypes.data, e.size)
    check("argot_hip_surrogate_trust_error")
    return e.reshape(trust_n, trust_n, trust_n)


def get_baseline():
    """(coefficients [20, 6], monomials in use: 0 or 20) of the polynomial
    baseline under the network."""
    c = np.zeros(abi.SURROGATE_NPOLY * NOUT, dtype=np.float32)
    n = lib().argot_hip_surrogate_get_baseline(c.ctypes.data)
    check("argot_hip_surrogate_get_baseline")
    return c.reshape(abi.SURROGATE_NPOLY, NOUT), n


def set_baseline(coef):
    """Set the baseline (coefficients [20, 6]), or remove it (None)."""
    if coef is None:
        lib().argot_hip_surrogate_set_baseline(None, 0)
    else:
        c = np.ascontiguousarray(coef, dtype=np.float32).reshape(-1)
        lib().argot_hip_surrogate_set_baseline(c.ctypes.data, abi.SURROGATE_NPOLY)
    check("argot_hip_surrogate_set_baseline")


def seeds_from_rays(tau, run, csect):
    """log10 threshold optical depths [n,3] of rays with walk optical depths
    tau [n,3] (the optical_depth_* sums of calc_optical_depth, in code
    units): tau * lunit * denstonh * sigma_threshold, the network's input
    scale.  Rays with a zero optical depth in some species are dropped."""
    cs = np.array([[c.csect_HI, c.csect_HeI, c.csect_HeII] for c in csect[:abi.NGRID_NU]],
                  dtype=np.float32)
    sig = np.array([cs[np.nonzero(cs[:, s])[0][0], s] for s in range(3)], dtype=np.float64)
    t = np.asarray(tau, dtype=np.float64) * (run.lunit * run.denstonh) * sig
    keep = np.all(t > 0.0, axis=1)
    return np.log10(t[keep]).astype(np.float32)
