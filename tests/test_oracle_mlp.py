"""Pins of the NumPy MLP reference (CPU): bf16 rounding on known values and
float64 gradients against central finite differences."""

import numpy as np
import numpy.testing as npt

from oracle import mlp_ref as R


def test_bf16_round_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -3.140625, 1e-3])
    got = R.bf16_round(x)
    # 1+2^-8 is a tie between 1 and 1+2^-7 -> even (1.0); 1+3*2^-8 ties -> 1+2^-6
    npt.assert_array_equal(got[:4], [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6])
    assert got[4] == -3.140625
    assert abs(got[5] - 1e-3) / 1e-3 < 2 ** -8


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    offs, o4, ob4, n = R.layout()
    p = np.zeros(n)
    for l, (ow, ob) in enumerate(offs):
        fan_in = R.N_IN if l == 0 else R.N_H
        W = rng.normal(0, np.sqrt(2.0 / fan_in), (R.N_H, R.N_PAD if l == 0 else R.N_H))
        if l == 0:
            W[:, R.N_IN:] = 0
        p[ow:ob] = W.ravel()
        p[ob:ob + R.N_H] = rng.normal(0, 0.1, R.N_H)
    p[o4:ob4] = rng.normal(0, np.sqrt(1.0 / R.N_H), R.N_H)
    p[ob4] = 0.3
    P = rng.uniform(R.N_IN * [0.0], R.N_IN * [1.0], (4, R.N_IN))
    X = R.features(P, np.zeros(R.N_IN), np.ones(R.N_IN))
    t = rng.normal(0, 1, 4)
    loss, g = R.loss_and_grad(p, X, t)
    # coordinates from every block: W0 (real inputs), b0, W2, b3, w4, b4
    picks = [offs[0][0] + 3, offs[0][0] + R.N_PAD * 7 + 5, offs[0][1] + 11,
             offs[2][0] + 1234, offs[3][1] + 77, o4 + 9, ob4]
    h = 1e-6
    for k in picks:
        e = np.zeros(n)
        e[k] = h
        fd = (R.loss_and_grad(p + e, X, t)[0] - R.loss_and_grad(p - e, X, t)[0]) / (2 * h)
        npt.assert_allclose(g[k], fd, rtol=1e-5, atol=1e-9)
    # padded input columns get exactly zero gradient
    assert np.all(g[offs[0][0]:offs[0][1]].reshape(R.N_H, R.N_PAD)[:, R.N_IN:] == 0)
