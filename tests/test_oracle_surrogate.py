"""The surrogate's oracle (CPU): the splitmix64 generator against its
published outputs, bf16 rounding against torch's, the exact kernels against
the restated reference rate loop in the thin-cell limit, and the FP64
gradient against central finite differences."""
import numpy as np
import pytest

import mlp_ref as M
from helpers import cell_centre
from test_oracle_rate import one_cell_problem

HNUL = 6.626e-27 * 3.28e15


def test_splitmix64_known_answers():
    # splitmix64 with state 1234567: the first five outputs of the
    # published generator (state += 0x9e3779b97f4a7c15 before each output).
    golden = 0x9E3779B97F4A7C15
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    states = [(1234567 + k * golden) & 0xFFFFFFFFFFFFFFFF for k in range(5)]
    got = M.mix64(np.array(states, dtype=np.uint64))
    assert [int(v) for v in got] == want


def test_uniform01_range_and_moments():
    u = M.uniform01(7, np.arange(1 << 16))
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.002
    x = M.sample_log_tau(7, 1 << 14, -6.0, 4.0)
    assert x.dtype == np.float32 and x.min() >= -6.0 and x.max() <= 4.0


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=4096).astype(np.float32),
                        np.float32([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0,
                                    1e-30, 3e38])])
    want = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    np.testing.assert_array_equal(M.bf16_round(x), want)
    assert M.bf16_round(np.float32([1.0 + 2 ** -8]))[0] == 1.0          # tie to even
    assert M.bf16_round(np.float32([1.0 + 3 * 2 ** -8]))[0] == 1.015625  # tie to even


def test_kernels_match_reference_rate_loop_in_thin_cell():
    """A ray into a tenuous target cell: the restated reference loop
    (calc_photoion_rate.c:285-387, expm1 form) gives Gamma_s = geo Q F_s and
    Heat_s = geo Q G_s hplanck nu_L with the column up to the cell."""
    P, w = one_cell_problem(n=64, dens=1e-8, fHI=1.0, fHeI=1.0, fHeII=1.0)
    cs = P.cs.reshape(-1, 3).astype(np.float64)
    nu = P.nu.astype(np.float32).astype(np.float64)
    spec = 1e46 * np.exp(-np.arange(256) / 60.0)
    sth = M.sigma_threshold(cs)
    col = w.lunit * w.denstonh
    tgt, src = (40, 10, 22), (0.1, 0.2, 0.3)
    c = np.array([cell_centre(v, 64) for v in tgt], np.float32)
    d = c - np.array(src, np.float32)
    r2 = float(np.float32(np.sqrt(np.float64(np.sum(d ** 2)))) ** 2)
    delta = float(P.delta_length(*d))
    geo = 1.0 / (4 * np.pi * w.lunit ** 2 * (r2 + delta ** 2 / 12))
    for lt in ([-3.0, -3.0, -3.0], [0.5, -1.0, -2.0], [1.3, 0.7, 0.2], [-1.0, 1.5, 0.8]):
        tau_ray = 10.0 ** np.array(lt) / sth / col      # code-unit columns
        out = P.ray_rate(np.array(src, np.float32), spec, *tgt, tau_ray, mode=1)
        k = 10.0 ** M.kernels(nu, cs, spec, np.array([lt]))[0]
        Q = spec.sum()
        np.testing.assert_allclose(out[[0, 2, 3]] / (geo * Q), k[:3], rtol=2e-6)
        np.testing.assert_allclose(out[[1, 4, 5]] / (geo * Q * HNUL), k[3:], rtol=2e-6)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = M.quantize(rng.normal(scale=0.3, size=M.NPARAM).astype(np.float32)).astype(np.float64)
    norm = np.concatenate([[-1.0, 0.2], rng.normal(size=6), rng.uniform(0.5, 2.0, 6)])
    x = rng.uniform(-6, 4, (32, 3))
    y = norm[2:8] + norm[8:] * rng.normal(size=(32, 6))
    loss, g = M.loss_and_grad(P, norm, x, y)
    # forward and loss agree
    t = (y - norm[2:8]) / norm[8:]
    yn = (M.forward(P, norm, x) - norm[2:8]) / norm[8:]
    assert abs(np.mean((yn - t) ** 2) - loss) < 1e-12 * max(loss, 1)
    idx = np.concatenate([rng.choice(np.arange(a, b), 12, replace=False)
                          for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2),
                                       (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                                       (M.OFF_W3, M.OFF_B3))] + [np.arange(M.OFF_B3, M.NPARAM)])
    h = 1e-6
    for i in idx:
        Pp = P.copy(); Pp[i] += h
        Pm = P.copy(); Pm[i] -= h
        fd = (M.loss_and_grad(Pp, norm, x, y)[0] - M.loss_and_grad(Pm, norm, x, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_seeded_samples_pick_seeds_and_jitter_normally():
    """Zero jitter reproduces the chosen seed points exactly; the seed index
    is uniform over the seeds; with jitter the offsets are N(0, jitter^2)
    per axis (mean, standard deviation, and the fraction beyond 2 sigma,
    which a wrong Box-Muller constant would move); clamping holds."""
    rng = np.random.default_rng(3)
    seeds = rng.uniform(-2.0, 1.0, (7, 3)).astype(np.float32)
    x0 = M.sample_seeded(11, 20000, seeds, 0.0, -6.0, 4.0)
    match = (x0[:, None, :] == seeds[None, :, :]).all(2)
    assert match.any(1).all()
    counts = match.argmax(1)
    freq = np.bincount(counts, minlength=7) / len(counts)
    assert np.allclose(freq, 1 / 7, atol=0.015)
    jit = 0.25
    x = M.sample_seeded(11, 200000, seeds, jit, -6.0, 4.0).astype(np.float64)
    # the seed of each sample is the zero-jitter draw with the same counter
    x0 = M.sample_seeded(11, 200000, seeds, 0.0, -6.0, 4.0).astype(np.float64)
    d = (x - x0) / jit
    assert abs(d.mean()) < 0.01
    assert abs(d.std() - 1.0) < 0.01
    assert abs(np.mean(np.abs(d) > 2.0) - 0.0455) < 0.003
    xc = M.sample_seeded(11, 5000, seeds, 3.0, -1.0, 0.5)
    assert xc.min() >= -1.0 and xc.max() <= 0.5


def test_baseline_monomials_and_forward():
    """The 20 monomials are 1, z and the products of degree 2 and 3 in
    lexicographic order of their index tuples (checked against
    itertools), and the forward pass adds monomials(z) @ poly."""
    import itertools
    z = np.array([[2.0, 3.0, 5.0], [-1.5, 0.25, 7.0]])
    want = [np.ones(2)] + [np.prod(z[:, list(c)], axis=1) for d in (1, 2, 3)
                           for c in itertools.combinations_with_replacement(range(3), d)]
    assert np.array_equal(M.monomials(z), np.stack(want, 1))
    assert M.monomials(z)[0, 14] == 2.0 * 3.0 * 5.0          # z0 z1 z2
    rng = np.random.default_rng(5)
    P = rng.normal(scale=0.3, size=M.NPARAM)
    norm = np.concatenate([[-1.0, 0.2], rng.normal(size=6), rng.uniform(1, 2, 6)])
    poly = rng.normal(size=(M.NPOLY, M.NOUT))
    x = rng.uniform(-3, 2, (50, 3))
    diff = M.forward(P, norm, x, poly) - M.forward(P, norm, x)
    np.testing.assert_allclose(diff, M.monomials((x - norm[0]) * norm[1]) @ poly, rtol=1e-12)
