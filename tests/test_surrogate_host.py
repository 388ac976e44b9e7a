"""Host logic of the surrogate's labels (CPU only): the label meshes and the
interpolated read-out at (S0, v0)."""

import numpy as np
import numpy.testing as npt

from fd_adi_amd.surrogate import (label_grids, label_price, draw_params, feller_ratio,
                                  BOX_LO, BOX_HI, VAR_MAX)


def test_label_meshes_scale_with_strike_and_ignore_v0():
    P = draw_params(6, 4)
    S, V = label_grids(P, 64, 32)
    unit = S[0] / (100.0 * P[0, 7])
    for s, v, p in zip(S, V, P):
        K = 100.0 * p[7]
        npt.assert_allclose(s / K, unit, rtol=1e-14, atol=1e-14)   # K x one unit mesh
        assert np.any(s == K) and np.all(np.diff(s) > 0) and s[0] == 0.0
        npt.assert_array_equal(v, V[0])                            # one variance mesh
        assert v[0] == 0.0 and v[-1] == VAR_MAX and np.all(np.diff(v) > 0)


def test_read_out_exact_for_bicubics():
    P = draw_params(16, 5)
    S, V = label_grids(P, 64, 32)

    def f(s, v):
        return 3.0 - 0.2 * s + 0.004 * s ** 2 - 1e-5 * s ** 3 + v * (2.0 + 0.01 * s) \
            - 5.0 * v ** 2 + 9.0 * v ** 3 * (1.0 + 1e-4 * s ** 3)

    G = np.stack([f(s[:, None], v[None, :]) for s, v in zip(S, V)])
    got = label_price(G, np.stack(S), np.stack(V), 100.0, P[:, 4])
    npt.assert_allclose(got, f(100.0, P[:, 4]), rtol=1e-12)


def test_read_out_at_nodes_returns_the_node_value():
    s = np.linspace(0.0, 10.0, 11)
    v = np.linspace(0.0, 1.0, 6)
    G = np.random.default_rng(0).random((1, 11, 6))
    npt.assert_allclose(label_price(G, s[None], v[None], 4.0, 0.6), G[0, 4, 3], rtol=1e-13)


def test_feller_ratio():
    P = np.array([[2.0, 0.04, 0.4, -0.5, 0.04, 0.03, 1.0, 1.0]])
    npt.assert_allclose(feller_ratio(P), [2 * 2.0 * 0.04 / 0.16])
    assert np.all(BOX_LO < BOX_HI)


def test_calibrate_threshold_ties_and_empty():
    from fd_adi_amd.surrogate import calibrate_threshold
    # the two lowest scores are tied and one of them fails: nothing qualifies
    tau, frac = calibrate_threshold([1.0, 1.0, 2.0], [0.0, 1.0, 0.0], atol=0.5)
    assert tau == -np.inf and frac == 0.0
    # the cut stops below the first failure
    tau, frac = calibrate_threshold([4.0, 1.0, 3.0, 2.0], [0.0, 0.0, 1.0, 0.0], atol=0.5)
    assert tau == 2.0 and frac == 0.5
    # a tie straddling the cut moves it below the tied value
    tau, frac = calibrate_threshold([1.0, 2.0, 2.0, 3.0], [0.0, 0.0, 1.0, 0.0], atol=0.5)
    assert tau == 1.0 and frac == 0.25
    # a failure budget of 20%: one failure in ten admitted
    s = np.arange(1.0, 11.0)
    e = np.zeros(10)
    e[2] = 1.0
    assert calibrate_threshold(s, e, atol=0.5, max_fail=0.2) == (10.0, 1.0)
    assert calibrate_threshold([], [], atol=0.5) == (-np.inf, 0.0)


def test_calibration_spread_of_one_failing_row():
    """One failing row among N, scored 300th lowest: a resample holds it with
    probability 1 - (1 - 1/N)^N = 0.632, and then admits the rows below it
    (about 0.30 of the resample); otherwise nothing fails and everything is
    admitted.  So the median sits near 0.30 and the 95th percentile at 1."""
    from fd_adi_amd.surrogate import calibration_spread
    N = 1000
    s = np.arange(N, dtype=float)
    e = np.zeros(N)
    e[300] = 1.0
    q = calibration_spread(s, e, atol=0.5, n=400, seed=3)
    assert abs(q["0.5"] - 0.30) < 0.04
    assert q["0.95"] == 1.0
    assert 0.2 < q["0.05"] <= q["0.5"]
    # no failing row: every resample admits everything
    q = calibration_spread(s, np.zeros(N), atol=0.5, n=50)
    assert q == {"0.05": 1.0, "0.5": 1.0, "0.95": 1.0}
    assert calibration_spread([], [], n=5) == {"0.05": 0.0, "0.5": 0.0, "0.95": 0.0}


def test_gate_rejects_everything_without_a_threshold():
    from fd_adi_amd.surrogate import SurrogatePricer, Ensemble
    P = draw_params(8, 1)

    class One(object):
        members = [None]

        def predict(self, P):
            return np.zeros(len(P)), np.zeros(len(P))

    # a one-network ensemble has no spread: refused without an error model
    import pytest
    with pytest.raises(ValueError):
        SurrogatePricer(One(), 0.0)
    pr = SurrogatePricer(One(), -np.inf, error_model=lambda P, s: np.zeros(len(P)))
    assert not pr.gate(P, np.zeros(len(P))).any()
    assert isinstance(Ensemble([None, None]).members, list)


def test_fallback_batches_cover_the_rows_with_little_padding():
    """AdiLabeler._batches: full batches of the largest tier first, then the
    remainder in tiers that cover it; padding at most an eighth of the last
    tier, or at most the smallest tier when nothing smaller fits."""
    from fd_adi_amd.surrogate import AdiLabeler
    lab = AdiLabeler(tiers=(256, 512, 1024, 2048, 4096), tune=False)
    rng = np.random.default_rng(3)
    for m in list(range(1, 600)) + list(rng.integers(1, 60000, 400)):
        caps = lab._batches(int(m))
        assert all(c in lab.tiers for c in caps)
        assert sum(caps) >= m and sum(caps[:-1]) < m       # every batch holds real rows
        assert caps[:int(m) // 4096] == [4096] * (int(m) // 4096)
        waste = sum(caps) - m
        assert waste <= max(caps[-1] // 8, 255), (m, caps)
    assert lab._batches(14350) == [4096, 4096, 4096, 2048, 256]
