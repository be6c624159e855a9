"""Photon-conserving rates of the oracle (CPU), pinned by the optically thin
closed form, photon conservation in an opaque cell, the species split, the
in-cell chord's closed forms and the degenerate source-in-cell case."""
import numpy as np
import pytest

from helpers import cell_centre, oracle_problem, small_workload

HNUL = 6.626e-27 * 3.28e15


def one_cell_problem(n=64, dens=1.0, fHI=1.0, fHeI=0.0, fHeII=0.0):
    w = small_workload(n=n, nsrc=1, seed=1, lognormal=False)
    w.dens[:] = dens
    w.fHI[:] = fHI
    w.fHeI[:] = fHeI
    w.fHeII[:] = fHeII
    return oracle_problem(w), w


def test_delta_length_closed_forms():
    P, w = one_cell_problem(n=64)
    h = 1.0 / 64
    assert P.delta_length(0.3, 0.0, 0.0) == np.float32(h)
    assert P.delta_length(0.0, -0.2, 0.0) == np.float32(h)
    assert abs(P.delta_length(0.1, 0.1, 0.0) / (h * np.sqrt(2)) - 1) < 1e-6
    assert abs(P.delta_length(-0.1, 0.1, 0.1) / (h * np.sqrt(3)) - 1) < 1e-6
    rng = np.random.default_rng(0)
    for _ in range(100):
        d = rng.normal(size=3)
        expect = h * np.linalg.norm(d) / np.abs(d).max()
        assert abs(P.delta_length(*d.astype(np.float32)) / expect - 1) < 1e-5


@pytest.mark.parametrize("mode", [0, 1])
def test_optically_thin_limit(mode):
    """tau -> 0 and a tenuous target: Gamma_s = sum_nu N sigma_s /
    (4 pi lunit^2 (r^2 + delta^2/12)), the free-streaming rate with the
    shell-averaged 1/r^2 (V_sh = 4 pi (r^2 delta + delta^3/12) lunit^3)."""
    P, w = one_cell_problem(n=64, dens=1e-8, fHI=1.0, fHeI=1.0, fHeII=1.0)
    cs = P.cs.reshape(-1, 3).astype(np.float64)
    spec = P.src_spec.reshape(-1)[:256] * 0 + 1e46 * np.exp(-np.arange(256) / 80.0)
    nu = P.nu.astype(np.float32).astype(np.float64)
    for tgt, src in [((40, 10, 22), (0.1, 0.2, 0.3)), ((5, 5, 60), (0.9, 0.1, 0.5))]:
        c = np.array([cell_centre(v, 64) for v in tgt], np.float32)
        d = c.astype(np.float32) - np.array(src, np.float32)
        r2 = float(np.float32(np.sqrt(np.float64(np.sum(d.astype(np.float32) ** 2)))) ** 2)
        delta = float(P.delta_length(*d))
        geo = 1.0 / (4 * np.pi * w.lunit ** 2 * (r2 + delta ** 2 / 12))
        # tau at the cell centre = half the cell's own optical depth
        col = w.lunit * w.denstonh
        ntgt = np.array([1e-8, 1e-8 * w.helium_fact, 1e-8 * w.helium_fact]) * w.denstonh
        tau_half = 0.5 * ntgt * delta * w.lunit / col
        out = P.ray_rate(np.array(src, np.float32), spec, *tgt, tau_half, mode=mode)
        thin = np.array([np.sum(spec * cs[:, s]) for s in range(3)]) * geo
        heat = np.array([np.sum(spec * cs[:, s] * np.maximum(nu - t, 0) * (nu > t))
                         for s, t in enumerate((1.0, 1.8125, 4.0))]) * geo * HNUL
        tol = 1e-6 if mode == 1 else 1e-5
        np.testing.assert_allclose(out[[0, 2, 3]], thin, rtol=tol)
        np.testing.assert_allclose(out[[1, 4, 5]], heat, rtol=tol)


def test_opaque_cell_absorbs_every_photon():
    """Entry optical depth 0 and an opaque target (huge n_HI): every photon
    of the ray is absorbed in the shell, Gamma_HI n_HI V_sh = sum_nu N_nu."""
    P, w = one_cell_problem(n=64, dens=1e9, fHI=1.0)
    spec = 1e46 * np.exp(-np.arange(256) / 50.0)
    tgt = (30, 31, 32)
    src = np.array([0.02, 0.7, 0.4], np.float32)
    c = np.array([cell_centre(v, 64) for v in tgt], np.float32)
    d = c - src
    delta = float(P.delta_length(*d))
    dist = float(np.float32(np.sqrt(np.float64(np.sum(d * d)))))
    nHI = 1e9 * w.denstonh
    col = w.lunit * w.denstonh
    tau = np.array([0.5 * nHI * delta * w.lunit / col, 0, 0])   # entry depth 0
    out = P.ray_rate(src, spec, *tgt, tau, mode=1)
    vsh = 4 * np.pi / 3 * ((dist + delta / 2) ** 3 - (dist - delta / 2) ** 3) * w.lunit ** 3
    assert abs(out[0] * nHI * vsh / spec.sum() - 1) < 1e-9


def test_species_split_conserves_absorbed_photons():
    """sum_s Gamma_s n_s V_sh = sum_nu N_nu exp(-tau_nu)(1 - exp(-dtau_nu)):
    the absorbed photons are shared among HI, HeI, HeII in proportion to
    their optical depths in each bin."""
    w = small_workload(n=32, nsrc=1, seed=5)
    P = oracle_problem(w)
    cs = P.cs.reshape(-1, 3).astype(np.float64)
    rng = np.random.default_rng(1)
    spec = 1e46 * rng.uniform(0.1, 1.0, 256)
    for _ in range(20):
        tgt = tuple(int(v) for v in rng.integers(0, 32, 3))
        src = rng.uniform(0, 1, 3).astype(np.float32)
        tau = rng.uniform(0, 3e-2, 3)
        out = P.ray_rate(src, spec, *tgt, tau, mode=1)
        c = np.array([cell_centre(v, 32) for v in tgt], np.float32)
        d = c - src
        delta = float(P.delta_length(*d))
        dist = float(np.float32(np.sqrt(np.float64(np.sum(d * d)))))
        vsh = 4 * np.pi / 3 * ((dist + delta / 2) ** 3 - (dist - delta / 2) ** 3) * w.lunit ** 3
        i = tgt
        nH = float(w.dens[i]) * w.denstonh
        n = np.array([nH * float(w.fHI[i]), nH * w.helium_fact * float(w.fHeI[i]),
                      nH * w.helium_fact * float(w.fHeII[i])])
        dtau_s = n * delta * w.lunit
        tin = tau * w.lunit * w.denstonh - 0.5 * dtau_s
        tnu = cs @ tin
        dnu = cs @ dtau_s
        absorbed = np.sum(spec * np.exp(-tnu) * -np.expm1(-dnu))
        lhs = out[0] * n[0] * vsh + out[2] * n[1] * vsh + out[3] * n[2] * vsh
        assert abs(lhs / absorbed - 1) < 1e-12


def test_reference_one_minus_exp_loses_digits():
    """The reference forms 1 - exp(-dtau) directly in FP64
    (calc_photoion_rate.c:359); for dtau ~ 1e-12 that keeps ~4 digits, which
    the FP64 oracle (expm1) does not share."""
    P, w = one_cell_problem(n=64, dens=1e-12, fHI=1.0)
    spec = np.full(256, 1e46)
    src = np.array([0.1, 0.1, 0.1], np.float32)
    ref = P.ray_rate(src, spec, 40, 40, 40, np.zeros(3), mode=0)
    exact = P.ray_rate(src, spec, 40, 40, 40, np.zeros(3), mode=1)
    err = abs(ref[0] / exact[0] - 1)
    assert 1e-7 < err < 1e-2


def test_source_in_target_cell_gives_inf_and_nan():
    """A ray whose source sits at the target cell centre (dist = 0): the
    chord is ~delta_x/1e-31, the entry optical depth -inf-like, exp(-tau)
    overflows and the reference yields Gamma_HI = Heat = inf and NaN for
    the helium rates (0 * inf in bins below their edges)."""
    P, w = one_cell_problem(n=64, dens=1.0, fHI=0.999, fHeI=0.999, fHeII=0.001)
    spec = np.full(256, 1e46)
    tgt = (10, 20, 30)
    src = np.array([cell_centre(v, 64) for v in tgt], np.float32)
    out = P.ray_rate(src, spec, *tgt, np.zeros(3), mode=0)
    assert np.isinf(out[0]) and np.isinf(out[1])
    assert np.isnan(out[2]) and np.isnan(out[3])
    assert np.isinf(out[4]) and np.isinf(out[5])
