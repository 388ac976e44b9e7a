"""Pins of the oracle's bicubic read-out (oracle/label.py)."""

import numpy as np
import numpy.testing as npt

from oracle.label import bicubic_at, stencil_start, lagrange4


def _mesh(n, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.uniform(0.2, 1.5, n)) - 0.3


def test_exact_for_bicubics_on_nonuniform_meshes():
    rng = np.random.default_rng(1)
    s, v = _mesh(20, 2), _mesh(12, 3)
    C = rng.normal(0, 1, (4, 4))

    def f(x, y):
        return sum(C[a, b] * x ** a * y ** b for a in range(4) for b in range(4))

    G = f(s[:, None], v[None, :])
    for _ in range(50):
        x = rng.uniform(s[0], s[-1])
        y = rng.uniform(v[0], v[-1])
        npt.assert_allclose(bicubic_at(G, s, v, x, y), f(x, y), rtol=1e-9, atol=1e-9)


def test_not_exact_for_quartics():
    """A plausible slip (a 3-point or a 5-point stencil) would change which
    polynomials are reproduced: degree 4 is not."""
    s, v = _mesh(20, 4), _mesh(12, 5)
    G = (s ** 4)[:, None] + 0 * v[None, :]
    x = 0.5 * (s[7] + s[8])
    assert abs(bicubic_at(G, s, v, x, v[3]) - x ** 4) > 1e-6


def test_node_value_and_stencil():
    s, v = _mesh(10, 6), _mesh(8, 7)
    G = np.random.default_rng(8).random((10, 8))
    npt.assert_allclose(bicubic_at(G, s, v, s[5], v[2]), G[5, 2], rtol=1e-13)
    # two nodes on each side between nodes 4 and 5; clipped at the ends
    assert stencil_start(s, 0.5 * (s[4] + s[5])) == 3
    assert stencil_start(s, s[0]) == 0 and stencil_start(s, s[-1]) == 6
    npt.assert_allclose(lagrange4(2.0, np.array([0.0, 1.0, 3.0, 4.0])).sum(), 1.0)
