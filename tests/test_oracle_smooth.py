"""smooth_photoion_rate's oracle (CPU) against closed forms: a constant
field is kept, a single spike spreads as 30/56 and 1/56, a linear field is
kept in the interior, and a corner spike keeps 37/56 (eight clamped copies
of the corner in its own stencil, one of them the central term)."""
import numpy as np

import oracle as O


def field(n, fn):
    ix, iy, iz = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    v = fn(ix, iy, iz).astype(np.float32).reshape(-1)
    return np.repeat(v[:, None], 6, axis=1) * np.float32([1, 2, 3, 4, 5, 6])


def test_constant_field_is_kept():
    r = field(6, lambda x, y, z: np.full(x.shape, 3.7))
    np.testing.assert_allclose(O.smooth_rates(r, (6, 6, 6)), r, rtol=1e-6)


def test_spike_spreads_with_the_stencil_weights():
    n = 7
    r = field(n, lambda x, y, z: ((x == 3) & (y == 3) & (z == 3)) * 56.0)
    s = O.smooth_rates(r, (n, n, n)).reshape(n, n, n, 6)
    np.testing.assert_allclose(s[3, 3, 3, 0], 30.0, rtol=1e-6)
    for d in [(2, 3, 3), (4, 4, 4), (2, 2, 2), (3, 4, 2)]:
        np.testing.assert_allclose(s[d][0], 1.0, rtol=1e-6)
    assert s[1, 3, 3, 0] == 0.0 and s[3, 3, 5, 0] == 0.0
    np.testing.assert_allclose(s.sum(axis=(0, 1, 2)), r.sum(0), rtol=1e-6)


def test_linear_field_is_kept_in_the_interior():
    n = 8
    r = field(n, lambda x, y, z: 1.0 + 0.5 * x - 0.25 * y + 0.125 * z)
    s = O.smooth_rates(r, (n, n, n)).reshape(n, n, n, 6)
    np.testing.assert_allclose(s[1:-1, 1:-1, 1:-1], r.reshape(n, n, n, 6)[1:-1, 1:-1, 1:-1],
                               rtol=1e-6)


def test_corner_spike_counts_its_clamped_copies():
    n = 5
    r = field(n, lambda x, y, z: ((x == 0) & (y == 0) & (z == 0)) * 56.0)
    s = O.smooth_rates(r, (n, n, n)).reshape(n, n, n, 6)
    np.testing.assert_allclose(s[0, 0, 0, 0], 37.0, rtol=1e-6)
    # an edge neighbour sees the corner through four clamped copies
    np.testing.assert_allclose(s[1, 0, 0, 0], 4.0, rtol=1e-6)
