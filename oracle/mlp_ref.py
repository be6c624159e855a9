"""ORACLE (test infrastructure, not product code): FP64 numpy restatement of
the rate surrogate -- its sample generator, the exact unit-spectrum kernels
it learns, and the network's forward pass and gradient.  Only tests/ use it,
as the checker of the GPU training and inference kernels.

The kernels are the frequency loop of src/calc_photoion_rate.c:285-387 in
the limit of a thin target cell (the factor (1-exp(-dtau))/dtau -> 1):
  F_s(N) = sum_k S_k sigma_s(k) exp(-sum_s' N_s' sigma_s'(k)),
  G_s(N) = sum_k S_k sigma_s(k) (nu_k - nu_s) exp(-...),  nu_k > nu_s,
with nu_s = 1, 1.8125, 4 (calc_photoion_rate.c:343-385).  Inputs are log10
of the threshold optical depths N_s sigma_s,th (sigma_s,th = first nonzero
table value of species s).

Sample points come from the splitmix64 finaliser (Steele, Lea & Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014).
"""
import numpy as np

NOUT = 6
HIDDEN = 64
THRESH = np.array([1.0, 1.8125, 4.0])
OFF_W1 = 0
OFF_B1 = OFF_W1 + 3 * HIDDEN
OFF_W2 = OFF_B1 + HIDDEN
OFF_B2 = OFF_W2 + HIDDEN * HIDDEN
OFF_W3 = OFF_B2 + HIDDEN
OFF_B3 = OFF_W3 + HIDDEN * NOUT
NPARAM = OFF_B3 + NOUT

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(z):
    """splitmix64 output function of state z (uint64 array, wrapping)."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform01(seed, i):
    """Uniform doubles in [0, 1) of counter i under seed."""
    h = mix64(np.uint64(seed) ^ mix64(np.asarray(i, dtype=np.uint64)))
    return (h >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def sample_log_tau(seed, n, lo, hi):
    """The generator's first n inputs: log10 tau [n,3], rounded to FP32."""
    i = np.arange(n, dtype=np.uint64)
    x = np.empty((n, 3), dtype=np.float32)
    for s in range(3):
        x[:, s] = (lo + (hi - lo) * uniform01(seed, np.uint64(3) * i + np.uint64(s))).astype(
            np.float32)
    return x


def sample_seeded(seed, n, seeds, jitter, lo, hi):
    """The generator's first n inputs with seed points: seed point
    j = floor(u * nseed) (u uniform under seed ^ 0x5eed5eed), plus jitter
    times a Box-Muller normal per axis (Box & Muller, Ann. Math. Stat. 29,
    1958: sqrt(-2 ln(1 - u1)) cos(2 pi u2), counters 6i + 2s and 6i + 2s + 1),
    clamped to [lo, hi] and rounded to FP32."""
    seeds = np.asarray(seeds, dtype=np.float32).reshape(-1, 3)
    i = np.arange(n, dtype=np.uint64)
    j = np.minimum((uniform01(np.uint64(seed) ^ np.uint64(0x5EED5EED), i) * len(seeds))
                   .astype(np.int64), len(seeds) - 1)
    x = np.empty((n, 3), dtype=np.float32)
    for s in range(3):
        u1 = uniform01(seed, np.uint64(6) * i + np.uint64(2 * s))
        u2 = uniform01(seed, np.uint64(6) * i + np.uint64(2 * s + 1))
        z = np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * np.pi * u2)
        v = seeds[j, s].astype(np.float64) + float(np.float32(jitter)) * z
        x[:, s] = np.clip(v, lo, hi).astype(np.float32)
    return x


NPOLY = 20


def monomials(z):
    """The 20 monomials of degree <= 3 of z [n,3] in the baseline's order:
    1, z0, z1, z2, then the degree-2 and degree-3 products in
    lexicographic order of their (sorted) index tuples."""
    z = np.asarray(z, dtype=np.float64)
    cols = [np.ones(len(z))]
    for d in (1, 2, 3):
        for c in _combos(d):
            cols.append(np.prod(z[:, list(c)], axis=1))
    return np.stack(cols, 1)


def _combos(d):
    out = []
    def rec(start, acc):
        if len(acc) == d:
            out.append(tuple(acc))
            return
        for k in range(start, 3):
            rec(k, acc + [k])
    rec(0, [])
    return out


def sigma_threshold(cs):
    cs = np.asarray(cs, dtype=np.float64).reshape(-1, 3)
    return np.array([cs[np.nonzero(cs[:, s])[0][0], s] for s in range(3)])


def kernels(nu, cs, spec, log_tau):
    """Exact log10 kernels [n,6] (FP64) at log10 threshold optical depths."""
    cs = np.asarray(cs, dtype=np.float64).reshape(-1, 3)
    nu = np.asarray(nu, dtype=np.float64)
    S = np.asarray(spec, dtype=np.float64) / np.sum(spec)
    N = 10.0 ** np.asarray(log_tau, dtype=np.float64) / sigma_threshold(cs)
    e = np.exp(-(N @ cs.T)) * S                     # [n, ngrid]
    W = np.zeros((len(nu), NOUT))
    for s in range(3):
        W[:, s] = cs[:, s]
        W[:, 3 + s] = np.where(nu > THRESH[s], cs[:, s] * (nu - THRESH[s]), 0.0)
    return np.log10(e @ W)


def bf16_round(x):
    """Round FP32 values to the nearest bf16 (ties to even), as FP32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)
    return (u << np.uint64(16)).astype(np.uint32).view(np.float32)


def quantize(P):
    """Weights rounded to bf16, biases unchanged (FP32)."""
    P = np.asarray(P, dtype=np.float32).copy()
    for a, b in ((OFF_W1, OFF_B1), (OFF_W2, OFF_B2), (OFF_W3, OFF_B3)):
        P[a:b] = bf16_round(P[a:b])
    return P


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return (P[OFF_W1:OFF_B1].reshape(3, HIDDEN), P[OFF_B1:OFF_W2],
            P[OFF_W2:OFF_B2].reshape(HIDDEN, HIDDEN), P[OFF_B2:OFF_W3],
            P[OFF_W3:OFF_B3].reshape(HIDDEN, NOUT), P[OFF_B3:NPARAM])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(Pq, norm, log_tau, poly=None):
    """log10 kernels [n,6] of the network with parameters Pq (already
    quantized) and normalisation norm = (xc, xs, ymean[6], ystd[6]), plus
    the polynomial baseline poly [20, 6] of the normalised input if given."""
    W1, b1, W2, b2, W3, b3 = unpack(Pq)
    norm = np.asarray(norm, dtype=np.float64)
    x = (np.asarray(log_tau, dtype=np.float64) - norm[0]) * norm[1]
    h1 = _silu(x @ W1 + b1)
    h2 = _silu(h1 @ W2 + b2)
    y = norm[2:2 + NOUT] + norm[2 + NOUT:] * (h2 @ W3 + b3)
    if poly is not None:
        y = y + monomials(x) @ np.asarray(poly, dtype=np.float64).reshape(NPOLY, NOUT)
    return y


def loss_and_grad(Pq, norm, log_tau, log_kernel):
    """Normalised mean squared error and its gradient with respect to the
    (quantized) parameters -- the straight-through gradient that training
    applies to the FP32 master copy."""
    W1, b1, W2, b2, W3, b3 = unpack(Pq)
    norm = np.asarray(norm, dtype=np.float64)
    x = (np.asarray(log_tau, dtype=np.float64) - norm[0]) * norm[1]
    t = (np.asarray(log_kernel, dtype=np.float64) - norm[2:2 + NOUT]) / norm[2 + NOUT:]
    z1 = x @ W1 + b1
    h1 = _silu(z1)
    z2 = h1 @ W2 + b2
    h2 = _silu(z2)
    y = h2 @ W3 + b3
    n = len(x)
    e = y - t
    loss = np.sum(e * e) / (n * NOUT)
    dy = 2.0 * e / (n * NOUT)
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = (h2.T @ dy).reshape(-1)
    g[OFF_B3:NPARAM] = dy.sum(0)
    dz2 = (dy @ W3.T) * _silu_grad(z2)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2).reshape(-1)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * _silu_grad(z1)
    g[OFF_W1:OFF_B1] = (x.T @ dz1).reshape(-1)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, g
