"""Device-resident serving (csrc/serve.hip): the gate's regression trees on
the device against scikit-learn, the gate's ordered list of rejected rows
against NumPy, the device-side engine update against the host one, and a
fitted surrogate served from device memory against its host-gated answers."""

import ctypes

import numpy as np
import numpy.testing as npt
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S(gpu):
    from fd_adi_amd import surrogate
    return surrogate


def test_forest_matches_scikit_learn(S):
    from sklearn.ensemble import HistGradientBoostingRegressor
    rng = np.random.default_rng(5)
    X = rng.standard_normal((20000, 10))
    y = np.sin(X[:, 0]) * X[:, 3] + np.abs(X[:, 7]) - 0.3 * X[:, 9] ** 2
    model = HistGradientBoostingRegressor(max_iter=120, max_leaf_nodes=63,
                                          random_state=0).fit(X, y)
    f = S.Forest(model, 10)
    Xq = rng.standard_normal((50000, 10)) * 1.3
    # the same summation order (baseline, then tree by tree): bit-equal
    npt.assert_array_equal(f.predict(Xq), model.predict(Xq))


def test_gate_lists_rejected_rows_in_order(S):
    from fd_adi_amd import _lib
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    rng = np.random.default_rng(6)
    for n in (1, 255, 256, 257, 70001, 300000):
        score = rng.standard_normal(n)
        inside = (rng.random(n) < 0.9).astype(np.int32)
        tau = 0.3
        d_score = SizedArrayPtr(score)
        d_inside = S.DeviceInts(inside)
        d_ok = S.DeviceInts(np.zeros(n, np.int32))
        d_idx = S.DeviceInts(np.full(n, -1, np.int32))
        work = SizedArrayPtr().alloc(n // 256 + 4)
        nrej = _lib.c_i32()
        _lib.call("fdadi_serve_gate", d_score.p, d_inside.p, n, tau, d_ok.p, d_idx.p, work.p,
                  ctypes.byref(nrej), _lib.default_stream())
        ok = (inside != 0) & (score <= tau)
        npt.assert_array_equal(d_ok.to_numpy().astype(bool), ok)
        assert nrej.value == (~ok).sum()
        npt.assert_array_equal(d_idx.to_numpy()[:nrej.value], np.nonzero(~ok)[0])


def test_device_update_matches_host_update(S):
    """Labels with the batch inputs built on the device (fdadi_hb_update_dev)
    equal those with them built on the host (fdadi_hb_update): the same
    meshes and parameters, bit for bit, across tiers and a tail."""
    lab = S.AdiLabeler(tiers=(64, 256), tune=False)
    P = S.draw_params(256 + 64 + 17, 41)
    dev = lab.price(P)
    host = lab.price_host_batches(P)
    npt.assert_array_equal(dev, host)


def test_device_serving_matches_host_gate(S):
    """A fitted surrogate served from device memory: the admitted flags are
    the host gate's (box, scikit-learn's trees on the host features -- the
    parameters and the ensemble's statistics -- the calibrated threshold),
    the admitted prices the ensemble's, the rejected ones the FP64 ADI
    labels."""
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    P = S.draw_params(8192, 406)
    y = S.adi_prices(P)
    # a small, short fit: a loose tolerance (1e-2) makes the gate admit part of
    # the in-box queries and reject the rest, so that both routes are compared
    pricer, info = S.fit_surrogate(P, y, members=2, steps=1000, batch=2048, max_fail=0.01,
                                   atol=1e-2)
    assert pricer.device_ready()
    Q = S.draw_params(3000, 407, outside_frac=0.2)
    st = pricer.ensemble.stats(Q)
    mean, score = st.mean, pricer.error_model(Q, st)
    want_ok = pricer.gate(Q, score)
    n = Q.shape[0]
    Qd = SizedArrayPtr(np.ascontiguousarray(Q))
    out, ok = SizedArrayPtr().alloc(n), S.DeviceInts(np.zeros(n, np.int32))
    pricer.price_device(Qd, n, out, ok)
    got, got_ok = out.to_numpy()[:n], ok.to_numpy().astype(bool)
    npt.assert_array_equal(got_ok, want_ok)
    assert 0 < got_ok.sum() < n and pricer.last_fallback == (~got_ok).sum()
    npt.assert_array_equal(got[got_ok], mean[got_ok])
    npt.assert_array_equal(got[~got_ok], S.adi_prices(Q[~got_ok]))
