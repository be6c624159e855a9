"""Seeded synthetic inputs for the point-source radiation step.

Every generator is deterministic in its seed (numpy PCG64).  Units follow the
reference's cosmological-comparison setup: box [0,1)^3 in code units with
lunit = 6.6 kpc, code density 1 == n_H = 1e-3 cm^-3 (denstonh = 1e-3).

Workloads (sizes as the benchmark names them):
  uniform64   64^3 uniform H+He, 16 sources at random cell centres, Q = 5e48
  lognormal128 128^3 lognormal field, 1024 sources, dN/dL ~ L^-2 on [1e47,1e50]
  clumped256  256^3 lognormal field, 4096 sources in 8 Gaussian clumps
  art128      128^3 lognormal field with seeded ionized fractions (long rays)
"""
from dataclasses import dataclass, field

import numpy as np

MPC_CM = 3.086e24                  # constants.h:14
LUNIT_CM = 6.6e-3 * MPC_CM         # 6.6 kpc = 2.03676e22 cm
DENSTONH = 1.0e-3                  # n_H [cm^-3] per code density unit
MPROTON_G = 1.672623e-24           # constants.h:15
KBOLTZ = 1.38066e-16               # constants.h:19
YEAR_S = 3.156e7                   # constants.h:24
XHYDROGEN, YHELIUM = 0.755, 0.245  # prim_chem/chemistry.h:23-24
GAMM1_MONOATOMIC = 0.66666666666666666   # prim_chem/chemistry.h:18
TUNIT_S = 1.0e8 * YEAR_S           # 100 Myr


def chem_units(lunit=LUNIT_CM, denstonh=DENSTONH, tunit=TUNIT_S):
    """munit, tunit and uenetok of the reference's comparison set-up
    (src/test_IC/cmpproj_test2_IC.c:72-79): code density 1 == n_H = denstonh,
    u * uenetok * wmol = T [K]."""
    munit = denstonh * lunit * lunit * lunit * MPROTON_G / XHYDROGEN
    uenetok = GAMM1_MONOATOMIC * MPROTON_G / KBOLTZ * (lunit * lunit) / (tunit * tunit)
    return {"lunit": lunit, "munit": munit, "tunit": tunit, "uenetok": uenetok,
            "denstonh": denstonh}


def wmol(felec):
    """Mean molecular weight 1/((1 + x_e) X + Y/4) (chemistry.h:69), float32."""
    fe = np.asarray(felec, np.float64)
    return (1.0 / ((1.0 + fe) * XHYDROGEN + 0.25 * YHELIUM)).astype(np.float32)


def helium_fact(y_he):
    """n_He / n_H for helium mass fraction Y (chemistry.h:23-25 form)."""
    return y_he / (4.0 * (1.0 - y_he))


@dataclass
class Workload:
    name: str
    n: int                       # mesh cells per side
    dens: np.ndarray             # float32 [n,n,n], z fastest
    fHI: np.ndarray
    fHII: np.ndarray
    fHeI: np.ndarray
    fHeII: np.ndarray
    fHeIII: np.ndarray
    src_pos: np.ndarray          # float32 [nsrc,3]
    src_rate: np.ndarray         # float64 [nsrc] total ionizing photons / s
    src_temp: np.ndarray         # float32 [nsrc] blackbody temperature [K]
    lunit: float = LUNIT_CM
    denstonh: float = DENSTONH
    helium_fact: float = helium_fact(0.24)
    temperature: float = 1.0e4
    meta: dict = field(default_factory=dict)

    @property
    def nsrc(self):
        return len(self.src_rate)


def _cell_centres(rng, n, count):
    idx = rng.integers(0, n, size=(count, 3))
    return ((idx.astype(np.float64) + 0.5) / n).astype(np.float32), idx


def gaussian_field(n, seed, slope=-3.0):
    """Periodic Gaussian random field with P(k) ~ k^slope, zero mean, unit variance."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n, n))
    fk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(n)[:, None, None]
    ky = np.fft.fftfreq(n)[None, :, None]
    kz = np.fft.rfftfreq(n)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    amp = k2 ** (slope / 4.0)            # sqrt(P) = k^(slope/2)
    amp[0, 0, 0] = 0.0
    g = np.fft.irfftn(fk * amp, s=(n, n, n), axes=(0, 1, 2))
    g -= g.mean()
    g /= g.std()
    return g


def lognormal_density(n, seed, sigma_ln=1.0):
    """Code density exp(sigma g - sigma^2/2): mean 1, i.e. <n_H> = 1e-3 cm^-3."""
    g = gaussian_field(n, seed)
    return np.exp(sigma_ln * g - 0.5 * sigma_ln ** 2).astype(np.float32)


def power_law_rates(rng, count, lmin=1.0e47, lmax=1.0e50):
    """dN/dL ~ L^-2 on [lmin, lmax] by inverse transform."""
    u = rng.random(count)
    return 1.0 / (1.0 / lmin - u * (1.0 / lmin - 1.0 / lmax))


def uniform64(seed=1, n=64, nsrc=16):
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    dens = np.ones(shape, np.float32)
    fHI = np.full(shape, 0.999, np.float32)
    fHII = np.full(shape, 0.001, np.float32)
    fHeI = np.full(shape, 0.999, np.float32)
    fHeII = np.full(shape, 0.001, np.float32)
    fHeIII = np.zeros(shape, np.float32)
    pos, _ = _cell_centres(rng, n, nsrc)
    return Workload("uniform64", n, dens, fHI, fHII, fHeI, fHeII, fHeIII, pos,
                    np.full(nsrc, 5.0e48), np.full(nsrc, 1.0e5, np.float32),
                    meta={"seed": seed, "nsrc": nsrc})


def _lognormal_fractions(rng, shape):
    fHI = rng.uniform(1e-4, 1.0, shape).astype(np.float32)
    fHeI = rng.uniform(1e-4, 1.0, shape).astype(np.float32)
    fHII = (1.0 - fHI).astype(np.float32)
    fHeII = (1.0 - fHeI).astype(np.float32)
    return fHI, fHII, fHeI, fHeII, np.zeros(shape, np.float32)


def lognormal128(seed=2, n=128, nsrc=1024):
    rng = np.random.default_rng(seed)
    dens = lognormal_density(n, seed + 1000)
    fHI, fHII, fHeI, fHeII, fHeIII = _lognormal_fractions(rng, dens.shape)
    pos, _ = _cell_centres(rng, n, nsrc)
    rates = power_law_rates(rng, nsrc)
    return Workload("lognormal128", n, dens, fHI, fHII, fHeI, fHeII, fHeIII, pos, rates,
                    np.full(nsrc, 1.0e5, np.float32), meta={"seed": seed, "nsrc": nsrc})


def clumped_sources(rng, n, nsrc, nclump=8, sigma_cells=8.0):
    """nclump uniform centres; members N(centre, sigma) per axis, clipped to the
    box and snapped to cell centres (so several sources may share a cell)."""
    centres = rng.uniform(0, n, size=(nclump, 3))
    member = rng.integers(0, nclump, size=nsrc)
    p = centres[member] + rng.normal(0.0, sigma_cells, size=(nsrc, 3))
    idx = np.clip(np.floor(p), 0, n - 1).astype(np.int64)
    return ((idx + 0.5) / n).astype(np.float32), idx


def clumped256(seed=3, n=256, nsrc=4096, nclump=8, sigma_cells=8.0):
    rng = np.random.default_rng(seed)
    dens = lognormal_density(n, seed + 1000)
    fHI, fHII, fHeI, fHeII, fHeIII = _lognormal_fractions(rng, dens.shape)
    pos, _ = clumped_sources(rng, n, nsrc, nclump, sigma_cells)
    rates = power_law_rates(rng, nsrc)
    return Workload("clumped256", n, dens, fHI, fHII, fHeI, fHeII, fHeIII, pos, rates,
                    np.full(nsrc, 1.0e5, np.float32),
                    meta={"seed": seed, "nsrc": nsrc, "nclump": nclump,
                          "sigma_cells": sigma_cells})


def ionized_fractions(rng, shape):
    """Seeded ionization state for the diffuse transfer: x_HII ~ U[0,1];
    helium a ~ U[0,1] ionized once, b ~ U[0,1] of that twice."""
    x = rng.uniform(0.0, 1.0, shape)
    a = rng.uniform(0.0, 1.0, shape)
    b = rng.uniform(0.0, 1.0, shape)
    fHII = x.astype(np.float32)
    fHI = (1.0 - x).astype(np.float32)
    fHeI = (1.0 - a).astype(np.float32)
    fHeII = (a * (1.0 - b)).astype(np.float32)
    fHeIII = (a * b).astype(np.float32)
    return fHI, fHII, fHeI, fHeII, fHeIII


def art128(seed=5, n=128):
    """Diffuse (ART) workload: the lognormal field of lognormal128's recipe,
    seeded ionized fractions, T = 1e4 K, no point sources."""
    rng = np.random.default_rng(seed)
    dens = lognormal_density(n, seed + 1000)
    fHI, fHII, fHeI, fHeII, fHeIII = ionized_fractions(rng, dens.shape)
    return Workload(f"art{n}", n, dens, fHI, fHII, fHeI, fHeII, fHeIII,
                    np.zeros((0, 3), np.float32), np.zeros(0), np.zeros(0, np.float32),
                    meta={"seed": seed})


WORKLOADS = {"uniform64": uniform64, "lognormal128": lognormal128, "clumped256": clumped256,
             "art128": art128}
