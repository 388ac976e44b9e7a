"""Surrogate path at the shapes of the surrogate configurations: labels on
256 x 128 meshes with nt = 200 Craig-Sneyd steps (training data), and
serving with the ADI fallback on the same meshes.

The oracle runs in parallel CPU processes (tests/oracle_pool.py); its
read-out at (S0, v0) is oracle/label.py's.  Bars: the label grid within
the grid criterion of tests/parity.py; the label itself within 1e-12
relative, or within 4 x the oracle's own move under its rounding-order
probes where those move it more (some parameter rows sit near the
reference scheme's stability limit on these meshes).
"""

import numpy as np
import numpy.testing as npt
import pytest

from tests.oracle_pool import oracle_labels
from tests.parity import grid_tolerance, record, FLOOR_FACTOR, PRICE_RTOL

pytestmark = pytest.mark.gpu

NS, NV, NT = 256, 128, 200


@pytest.fixture(scope="module")
def S(gpu):
    from fd_adi_amd import surrogate
    return surrogate


@pytest.fixture(scope="module")
def labeler(S):
    return S.AdiLabeler(NS, NV, NT, tiers=(16, 64))


def label_bar(y, want, lfloor):
    return max(PRICE_RTOL * abs(want), FLOOR_FACTOR * lfloor)


def test_config3_labels_match_oracle(S, labeler):
    """16 parameter rows drawn over the training box: the device labels and
    the value grids behind them against the oracle's Craig-Sneyd solve on the
    same meshes; the device read-out against the oracle's read-out of the
    device grid."""
    from oracle.label import bicubic_at
    P = S.draw_params(16, 301)
    y = labeler.price(P)
    G = labeler.grids(P)
    spots, vars_ = S.label_grids(P, NS, NV)
    W, want, floors, lfloors = oracle_labels(P, spots, vars_, NT)
    rels = []
    for k in range(len(P)):
        tol, floor = grid_tolerance(W[k], floors[k])
        err = float(np.max(np.abs(G[k] - W[k])))
        rel = abs(y[k] - want[k]) / abs(want[k])
        rels.append(rel)
        print("row %2d: grid err %.2e (tol %.2e)  label %.10f rel %.2e (label floor %.2e)"
              % (k, err, tol, y[k], rel, lfloors[k]))
        record(err=err, floor=floor, tol=tol, ratio=err / floor if floor else None,
               price_rel=rel, case="label row %d" % k)
        assert err <= tol, (k, err, tol)
        assert abs(y[k] - want[k]) <= label_bar(y[k], want[k], lfloors[k]), (k, rel)
        # the device read-out of the device grid is the oracle's read-out of it
        host = bicubic_at(G[k], spots[k], vars_[k], 100.0, P[k, 4])
        assert abs(y[k] - host) <= 1e-13 * max(1.0, abs(host)), (k, y[k], host)
    print("label rel errors: median %.2e max %.2e" % (np.median(rels), np.max(rels)))
    assert np.median(rels) <= PRICE_RTOL


def test_labeler_batches_tails_and_graph_reuse(S, labeler):
    """Rows beyond one batch, a short tail, and the same rows priced twice
    (the engine, its graph and its buffers reused) give the same labels; one
    row at a time (the smallest tier, whose autotuned launch shape may chunk
    the line solves differently) agrees to rounding."""
    P = S.draw_params(37, 302)
    a = labeler.price(P)
    b = labeler.price(P[::-1])[::-1]
    c = np.array([labeler.price(P[k:k + 1])[0] for k in (0, 17, 36)])
    npt.assert_array_equal(a, b)
    npt.assert_allclose(a[[0, 17, 36]], c, rtol=1e-13, atol=0)


def test_labels_reproduce_bit_for_bit_at_one_launch_shape(S):
    """Fresh labellers (new engines, graphs and buffers) with the same launch
    shape give bit-identical labels, and per-option shapes that resolve to
    the same tiles (every row here is 256 x 128) give the same bits: the
    training data depends on the launch shape only, not on the run."""
    P = S.draw_params(64, 11)
    out = []
    for shape in [(256, 16), (256, 16), (-1, 16)]:
        lab = S.AdiLabeler(NS, NV, NT, tiers=(64,), tune=False)
        hb, _ = lab._engine(64)
        hb.set_shape(0, *shape)
        hb.set_shape(1, 8)
        hb.capture(NT)
        out.append(lab.price(P))
    assert np.all(np.isfinite(out[0]))
    npt.assert_array_equal(out[0], out[1])
    npt.assert_array_equal(out[0], out[2])


def test_network_gains_over_the_semi_analytic_price(S, labeler):
    """What the trained network adds: on held-out rows, the semi-analytic
    price plus the network's correction is closer to the FP64 ADI label than
    the semi-analytic price alone (the method without the network), in RMSE
    and in the share of rows within 1e-4 and 1e-3."""
    P = S.draw_params(16384 + 2048, 305)
    y = S.adi_prices(P, NS, NV, NT)
    base = S.cf_price(P)
    ok = S.within_call_bounds(P, y) & (np.abs(y - base) < 0.05)
    P, y, base = P[ok], y[ok], base[ok]
    tr, te = slice(0, 16384 - 512), slice(16384 - 512, None)
    ens, hist = S.train_ensemble(P[tr], y[tr], members=1, steps=3000, batch=4096,
                                 schedule="cosine")
    pred, _ = ens.predict(P[te])
    e_net, e_base = np.abs(pred - y[te]), np.abs(base[te] - y[te])
    frac = lambda e, a: float(np.mean(e <= a))
    rmse = lambda e: float(np.sqrt(np.mean(e ** 2)))
    print("held-out %d rows: rmse net %.2e base %.2e; within 1e-4 net %.3f base %.3f; "
          "within 1e-3 net %.3f base %.3f" % (e_net.size, rmse(e_net), rmse(e_base),
                                               frac(e_net, 1e-4), frac(e_base, 1e-4),
                                               frac(e_net, 1e-3), frac(e_base, 1e-3)))
    assert rmse(e_net) < 0.7 * rmse(e_base)
    assert frac(e_net, 1e-4) > frac(e_base, 1e-4) + 0.3
    assert frac(e_net, 1e-3) > frac(e_base, 1e-3)


def test_config4_serving_routes_to_oracle_checked_adi(S, labeler):
    """Serving on the 256 x 128 fallback meshes: queries the gate accepts get
    the semi-analytic price plus the ensemble's correction; the others (every
    query outside the training box among them) get the FP64 ADI label, here
    checked against the oracle's, not against the labeller itself."""
    P = S.draw_params(2048, 303)
    y = S.adi_prices(P, NS, NV, NT)
    ok = S.within_call_bounds(P, y) & (np.abs(y - S.cf_price(P)) < 0.05)
    ens, _ = S.train_ensemble(P[ok], y[ok], members=2, steps=200, batch=1024)
    Q = S.draw_params(48, 304, outside_frac=0.25)
    mean, spread = ens.predict(Q)
    tau = float(np.median(spread))
    pricer = S.SurrogatePricer(ens, tau, adi=labeler.price)
    price, accepted = pricer.price(Q)
    inside = np.all((Q >= S.BOX_LO) & (Q <= S.BOX_HI), axis=1)
    npt.assert_array_equal(accepted, inside & (spread <= tau))
    assert 0 < accepted.sum() < inside.sum() and (~inside).sum() >= 10
    npt.assert_array_equal(price[accepted], mean[accepted])
    spots, vars_ = S.label_grids(Q, NS, NV)
    W, want, floors, lfloors = oracle_labels(Q, spots, vars_, NT)
    for k in np.nonzero(~accepted)[0]:
        assert abs(price[k] - want[k]) <= label_bar(price[k], want[k], lfloors[k]), \
            (k, price[k], want[k], lfloors[k])
    err_acc = np.abs(price[accepted] - want[accepted])
    print("fallback %d of %d (outside the box %d); accepted: max |err| vs oracle %.2e"
          % ((~accepted).sum(), len(Q), (~inside).sum(), err_acc.max()))


def test_fitted_surrogate_gate_and_routing(S):
    """fit_surrogate end to end on a small label set: the error model and
    the calibrated threshold route the queries -- admitted ones lie inside
    the box and have score <= tau and get the surrogate's price, every other
    one (all those outside the box among them) gets the FP64 ADI label."""
    P = S.draw_params(8192, 306)
    y = S.adi_prices(P, NS, NV, NT)
    pricer, info = S.fit_surrogate(P, y, members=2, steps=300, batch=2048, max_fail=0.01)
    print(info)
    assert info["error_model"] == "trees" and info["calib_rows"] > 600
    Q = S.draw_params(512, 307, outside_frac=0.2)
    st = pricer.ensemble.stats(Q)
    mean, score = st.mean, pricer.error_model(Q, st)
    price, accepted = pricer.price(Q)
    inside = np.all((Q >= S.BOX_LO) & (Q <= S.BOX_HI), axis=1)
    npt.assert_array_equal(accepted, inside & (score <= pricer.tau))
    assert not accepted[~inside].any() and pricer.last_fallback == (~accepted).sum()
    npt.assert_array_equal(price[accepted], mean[accepted])
    npt.assert_array_equal(price[~accepted], S.adi_prices(Q[~accepted], NS, NV, NT))
