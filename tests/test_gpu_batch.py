"""Batched engine: uneven grids in one launch per stage, per-option parity
with the oracle, graph replay equal to eager stepping, and full-size
(2048 x 1024) properties."""

import numpy as np
import numpy.testing as npt
import pytest

pytestmark = pytest.mark.gpu

from oracle.engine import HestonOracle, HestonParams
from tests.parity import grid_tolerance, alternates, record, PRICE_RTOL, FLOOR_FACTOR
from fd_adi_amd.utils import sinh_space, exponential_space


@pytest.fixture(scope="module")
def fd(gpu):
    import fd_adi_amd
    return fd_adi_amd


def draw(seed, n, nS_range=(16, 64), nv_range=(8, 40)):
    rng = np.random.default_rng(seed)
    P, S, Vs, dts, O = [], [], [], [], []
    for k in range(n):
        kappa, theta = rng.uniform(0.5, 5), rng.uniform(0.01, 0.16)
        sigma, rho = rng.uniform(0.1, 1.0), rng.uniform(-0.9, 0.0)
        r, T, K = rng.uniform(0, 0.08), rng.uniform(0.1, 3), 100 * rng.uniform(0.7, 1.3)
        nS, nv = int(rng.integers(*nS_range)), int(rng.integers(*nv_range))
        spots = sinh_space(K, 5 * K, 7.0, nS)
        vars_ = exponential_space(0, theta, 5 * theta, 4.0, nv)
        nt = 20
        P.append(dict(r=r, kappa=kappa, theta=theta, sigma=sigma, rho=rho))
        S.append(spots)
        Vs.append(vars_)
        dts.append(T / nt)
        O.append(HestonOracle(HestonParams(spot=100, strike=K, r=r, v0=theta, T=T, kappa=kappa,
                                           theta=theta, sigma=sigma, rho=rho), spots, vars_))
    return P, S, Vs, dts, O


@pytest.mark.parametrize("scheme", ["craigsneyd", "hundsdorferverwer", "douglas"])
def test_batch_uneven_grids_match_oracle(fd, scheme):
    P, S, Vs, dts, O = draw(7, 6)
    hb = fd.HestonBatch(P, S, Vs, dts)
    hb.set_state([o.payoff() for o in O])
    hb.configure(scheme, 0.5)
    hb.step(20)
    got = hb.get_state()
    for k, o in enumerate(O):
        want = getattr(o, "solve_" + scheme)(20, dts[k], o.payoff(), theta=0.5)
        alt = [getattr(a, "solve_" + scheme)(20, dts[k], o.payoff(), theta=0.5)
               for a in alternates(o)]
        tol, floor = grid_tolerance(want, alt)
        err = np.max(np.abs(got[k] - want))
        rel = abs(got[k][o.idx] - want[o.idx]) / abs(want[o.idx])
        print(k, got[k].shape, "err %.3e tol %.3e floor %.3e price rel %.2e" % (err, tol, floor, rel))
        record(err=float(err), floor=floor, tol=tol, ratio=float(err) / floor if floor else None,
               price_rel=float(rel), case="option %d %s" % (k, got[k].shape))
        assert err <= tol, (k, err, tol)
        assert rel <= PRICE_RTOL, (k, rel)


def test_graph_replay_equals_eager(fd):
    P, S, Vs, dts, O = draw(11, 4)
    init = [o.payoff() for o in O]
    eager = fd.HestonBatch(P, S, Vs, dts)
    eager.set_state(init)
    eager.configure("cs", 0.5)
    eager.step(4)  # warm-up launches (kernel attributes) outside capture
    eager.step(12)
    ref = eager.get_packed()
    g = fd.HestonBatch(P, S, Vs, dts)
    g.configure("cs", 0.5)
    g.set_state(init)
    g.step(4)
    g.capture(4)
    for rep in range(3):
        g.replay()
        g.sync()
        # eager work and synchronisation between replays
        _ = g.get_packed()
    npt.assert_array_equal(g.get_packed(), ref)


def test_whole_solve_graph_with_smoothing(fd):
    """fdadi_hb_capture_solve: one graph for a smoothed solve (4 implicit half
    steps, then HV at theta 0.6, each part with its refactorisation), replayed
    three times from the payoff with eager work between the replays: every
    replay equals the eager solve bit for bit, and the eager solve matches the
    oracle's smoothed solve."""
    P, S, Vs, dts, O = draw(13, 4)
    init = [o.payoff() for o in O]
    eager = fd.HestonBatch(P, S, Vs, dts)
    eager.set_state(init)
    eager.solve("hv", 20, theta=0.6, smoothing_steps=2)
    ref = eager.get_packed()
    got = eager.get_state()
    for k, o in enumerate(O):
        want = o.solve_smooth(20, dts[k], o.payoff(), smoothing_steps=2)
        alt = [a.solve_smooth(20, dts[k], o.payoff(), smoothing_steps=2) for a in alternates(o)]
        tol, floor = grid_tolerance(want, alt)
        assert np.max(np.abs(got[k] - want)) <= tol, k
    g = fd.HestonBatch(P, S, Vs, dts)
    g.set_state(init)
    g.solve("hv", 20, theta=0.6, smoothing_steps=2)  # kernel attributes set outside capture
    g.capture_solve("hv", 20, theta=0.6, smoothing_steps=2)
    for rep in range(3):
        g.set_state(init)
        g.replay()
        g.sync()
        npt.assert_array_equal(g.get_packed(), ref)
        # eager work between the replays of the same graph
        eager.set_state(init)
        eager.solve("douglas", 3, theta=0.5)
        eager.sync()
    with pytest.raises(ValueError):
        g.capture_solve("hv", 2, theta=0.6, smoothing_steps=2)
    # after a capture the engine steps eagerly as configured for the scheme
    fresh = fd.HestonBatch(P, S, Vs, dts)
    fresh.capture_solve("hv", 20, theta=0.6, smoothing_steps=2)
    fresh.set_state(init)
    fresh.step(5)
    plain = fd.HestonBatch(P, S, Vs, dts)
    plain.set_state(init)
    plain.configure("hv", 0.6)
    plain.step(5)
    npt.assert_array_equal(fresh.get_packed(), plain.get_packed())


def test_launch_shapes_agree(fd):
    P, S, Vs, dts, O = draw(3, 3, nS_range=(40, 100), nv_range=(32, 64))
    outs = []
    for ts, nl, tv in ((0, 0, 0), (256, 4, 16), (512, 4, 64), (256, 8, 16), (512, 8, 8),
                       (256, 16, 8), (512, 16, 32), (512, 32, 16), (-1, 4, -1), (-1, 16, 16),
                       (-1, 8, -1)):
        hb = fd.HestonBatch(P, S, Vs, dts)
        if ts:
            hb.set_shape(0, ts, nl)
            hb.set_shape(1, tv)
            assert hb.get_shape(0) == (ts, nl)
        hb.set_state([o.payoff() for o in O])
        hb.configure("hv", 0.5)
        hb.step(10)
        outs.append(hb.get_packed())
    scale = np.max(np.abs(outs[0]))
    for x in outs[1:]:
        # chunking changes the summation order of the line solves only
        assert np.max(np.abs(x - outs[0])) <= 1e-12 * scale


def test_per_option_shapes_match_oracle(fd):
    """Launch shapes chosen per option (threads = -1): options whose spot
    lines (40 .. 500 nodes) and variance lines (20 .. 170) need different
    shapes go to different launches of one stage.  Craig-Sneyd, 4- and
    16-column tiles: every option agrees with the one-shape launch to the
    chunking's rounding (1e-12 of the grid's scale), and with the oracle.
    (The grids are ones the scheme is stable on: with rho = -0.6, or a
    variance mesh much finer than the spot mesh, the reference scheme
    amplifies rounding differences as on the 2048 x 1024 grid.)"""
    P, S, Vs, dts, O = [], [], [], [], []
    rng = np.random.default_rng(14)
    for nS, nv in ((40, 60), (70, 20), (130, 90), (260, 40), (500, 120), (300, 33), (450, 170)):
        K, theta = 100 * rng.uniform(0.8, 1.2), rng.uniform(0.02, 0.1)
        p = dict(r=0.03, kappa=2.0, theta=theta, sigma=0.4, rho=-0.3)
        spots = sinh_space(K, 5 * K, 7.0, nS)
        vars_ = exponential_space(0, theta, 5 * theta, 4.0, nv)
        P.append(p)
        S.append(spots)
        Vs.append(vars_)
        dts.append(1.0 / 10)
        O.append(HestonOracle(HestonParams(spot=100, strike=K, r=p["r"], v0=theta, T=1.0,
                                           kappa=p["kappa"], theta=theta, sigma=p["sigma"],
                                           rho=p["rho"]), spots, vars_))
    want = [o.solve_craigsneyd(10, dts[k], o.payoff(), theta=0.5) for k, o in enumerate(O)]
    alts = [[a.solve_craigsneyd(10, dts[k], o.payoff(), theta=0.5) for a in alternates(o)]
            for k, o in enumerate(O)]
    one = fd.HestonBatch(P, S, Vs, dts)
    one.set_state([o.payoff() for o in O])
    one.configure("cs", 0.5)
    one.step(10)
    ref = one.get_state()
    for nl in (4, 16):
        hb = fd.HestonBatch(P, S, Vs, dts)
        hb.set_shape(0, -1, nl)
        hb.set_shape(1, -1)
        assert hb.get_shape(0) == (-1, nl) and hb.get_shape(1) == (-1, 0)
        hb.set_state([o.payoff() for o in O])
        hb.configure("cs", 0.5)
        assert hb.launches_per_step() > 4  # more than one launch per stage
        hb.step(10)
        got = hb.get_state()
        for k in range(len(O)):
            scale = np.max(np.abs(ref[k]))
            assert np.max(np.abs(got[k] - ref[k])) <= 1e-12 * scale, (nl, k)
            tol, floor = grid_tolerance(want[k], alts[k])
            err = float(np.max(np.abs(got[k] - want[k])))
            assert err <= tol, (nl, k, got[k].shape, err, tol)


def test_every_column_residue_every_shape_matches_oracle(fd):
    """Variance sizes 16..33 cover every residue modulo the tile widths (4, 8,
    16): the last tile of an option then overlaps the one before it, and the
    last column takes its j-2 neighbour from within its tile.  Craig-Sneyd,
    10 steps, every spot launch shape against the oracle."""
    rng = np.random.default_rng(11)
    P, S, Vs, dts, O = [], [], [], [], []
    for nv in range(16, 34):
        nS = int(rng.integers(20, 44))
        K, theta = 100 * rng.uniform(0.8, 1.2), rng.uniform(0.02, 0.1)
        p = dict(r=0.03, kappa=2.0, theta=theta, sigma=0.4, rho=-0.6)
        spots = sinh_space(K, 5 * K, 7.0, nS)
        vars_ = exponential_space(0, theta, 5 * theta, 4.0, nv)
        P.append(p)
        S.append(spots)
        Vs.append(vars_)
        dts.append(1.0 / 10)
        O.append(HestonOracle(HestonParams(spot=100, strike=K, r=p["r"], v0=theta, T=1.0,
                                           kappa=p["kappa"], theta=theta, sigma=p["sigma"],
                                           rho=p["rho"]), spots, vars_))
    want = [o.solve_craigsneyd(10, dts[k], o.payoff(), theta=0.5) for k, o in enumerate(O)]
    alts = [[a.solve_craigsneyd(10, dts[k], o.payoff(), theta=0.5) for a in alternates(o)]
            for k, o in enumerate(O)]
    for shape in ((0, 0), (64, 4), (128, 4), (256, 4), (256, 8), (512, 8), (256, 16),
                  (512, 16), (-1, 4), (-1, 16)):
        hb = fd.HestonBatch(P, S, Vs, dts)
        if shape[0]:
            hb.set_shape(0, *shape)
        if shape[0] < 0:
            hb.set_shape(1, -1)
        hb.set_state([o.payoff() for o in O])
        hb.configure("cs", 0.5)
        hb.step(10)
        got = hb.get_state()
        for k in range(len(O)):
            tol, floor = grid_tolerance(want[k], alts[k])
            err = float(np.max(np.abs(got[k] - want[k])))
            assert err <= tol, (shape, k, got[k].shape, err, tol)


def test_32_column_tiles_every_residue_match_oracle(fd):
    """32-column spot tiles (neighbour columns across the tile's two DPP rows
    by wave shifts): variance sizes 32..65 cover every residue modulo 32, so
    the last tile overlaps the one before it; Craig-Sneyd and
    Hundsdorfer-Verwer, 10 steps, against the oracle.  Fewer than 32
    variance nodes are refused."""
    rng = np.random.default_rng(12)
    P, S, Vs, dts, O = [], [], [], [], []
    for nv in range(32, 66):
        nS = int(rng.integers(20, 44))
        K, theta = 100 * rng.uniform(0.8, 1.2), rng.uniform(0.02, 0.1)
        p = dict(r=0.03, kappa=2.0, theta=theta, sigma=0.4, rho=-0.6)
        spots = sinh_space(K, 5 * K, 7.0, nS)
        vars_ = exponential_space(0, theta, 5 * theta, 4.0, nv)
        P.append(p)
        S.append(spots)
        Vs.append(vars_)
        dts.append(1.0 / 10)
        O.append(HestonOracle(HestonParams(spot=100, strike=K, r=p["r"], v0=theta, T=1.0,
                                           kappa=p["kappa"], theta=theta, sigma=p["sigma"],
                                           rho=p["rho"]), spots, vars_))
    for scheme, solve in (("cs", "solve_craigsneyd"), ("hv", "solve_hundsdorferverwer")):
        want = [getattr(o, solve)(10, dts[k], o.payoff(), theta=0.5) for k, o in enumerate(O)]
        alts = [[getattr(a, solve)(10, dts[k], o.payoff(), theta=0.5) for a in alternates(o)]
                for k, o in enumerate(O)]
        hb = fd.HestonBatch(P, S, Vs, dts)
        hb.set_shape(0, 512, 32)
        assert hb.get_shape(0) == (512, 32)
        hb.set_state([o.payoff() for o in O])
        hb.configure(scheme, 0.5)
        hb.step(10)
        got = hb.get_state()
        for k in range(len(O)):
            tol, floor = grid_tolerance(want[k], alts[k])
            err = float(np.max(np.abs(got[k] - want[k])))
            assert err <= tol, (scheme, k, got[k].shape, err, tol)
    small = fd.HestonBatch(P[:1], [S[0]], [Vs[0][:31]], dts[:1])
    with pytest.raises(Exception):
        small.set_shape(0, 512, 32)


def test_full_size_config1_steps(fd):
    """2048 x 1024, rho = -0.7, HV: the first 50 steps against the oracle
    (the oracle and its three rounding-order probes run in parallel CPU
    processes), then the Dirichlet row after 20 more."""
    from tests.oracle_pool import solve_many
    from tests.parity import PROBES
    spots = sinh_space(100, 500, 7.0, 2048)
    vars_ = exponential_space(0, 0.04, 0.2, 4.0, 1024)
    spots[np.argmin(np.abs(spots - 100))] = 100
    P = [dict(r=0.06, kappa=1.0, theta=0.04, sigma=0.2, rho=-0.7)]
    pf = dict(spot=100, strike=100, r=0.06, v0=0.04, T=1.0, kappa=1.0, theta=0.04, sigma=0.2,
              rho=-0.7)
    O = HestonOracle(HestonParams(**pf), spots, vars_)
    n, dt = 50, 1 / 2000.0
    hb = fd.HestonBatch(P, [spots], [vars_], [dt])
    U0 = O.payoff()
    hb.set_state([U0])
    hb.configure("hv", 0.5)
    hb.step(n)
    got = hb.get_state()[0]
    res = solve_many([(pf, spots, vars_, dict(kw), "solve_hundsdorferverwer", (n, dt, U0, 0.5))
                      for kw in ({},) + PROBES])
    want, alt = res[0], res[1:]
    tol, floor = grid_tolerance(want, alt)
    err = np.max(np.abs(got - want))
    rel = abs(got[O.idx] - want[O.idx]) / abs(want[O.idx])
    print("2048x1024 HV %d steps: err %.3e tol %.3e floor %.3e price rel %.2e"
          % (n, err, tol, floor, rel))
    record(err=float(err), floor=floor, tol=tol, ratio=float(err) / floor if floor else None,
           price_rel=float(rel), case="2048x1024 HV %d steps" % n)
    assert err <= tol and rel <= PRICE_RTOL
    hb.step(20)
    V = hb.get_state()[0]
    assert np.all(V[0] == 0) and np.all(np.isfinite(V))


def test_reference_instability_reproduced(fd):
    """On the 1024-point power-law variance mesh (first spacing 1.8e-13) the
    reference algorithm itself is unstable at the v = vmax boundary row: the
    oracle's max|V| grows ~12% per step after ~100 steps.  The GPU engine
    reproduces the oracle there (512 x 1024, HV, 150 steps)."""
    from fd_adi_amd.workloads import single
    o = single(512, 1024, 2000, -0.7)
    O = HestonOracle(HestonParams(spot=100, strike=100, r=0.06, v0=0.04, T=1.0, kappa=1.0,
                                  theta=0.04, sigma=0.2, rho=-0.7), o.spots, o.vars)
    hb = fd.HestonBatch([o.params], [o.spots], [o.vars], [o.dt])
    hb.set_state([o.payoff()])
    hb.configure("hv", 0.5)
    hb.step(150)
    V = hb.get_state()[0]
    W = O.solve_hundsdorferverwer(150, o.dt, o.payoff())
    iv, iw = np.unravel_index(np.argmax(np.abs(V)), V.shape), np.unravel_index(
        np.argmax(np.abs(W)), W.shape)
    print("GPU max|V| %.6e at %s, oracle %.6e at %s" % (np.abs(V).max(), iv, np.abs(W).max(), iw))
    assert np.abs(W).max() > 1e3            # the oracle has started to diverge
    assert iv == iw
    npt.assert_allclose(np.abs(V).max(), np.abs(W).max(), rtol=1e-4)
    npt.assert_allclose(V[o.idx], W[o.idx], rtol=1e-10)


def test_autotune_restores_state_and_keeps_results(fd):
    P, S, Vs, dts, O = draw(5, 3, nS_range=(40, 120), nv_range=(20, 80))
    init = [o.payoff() for o in O]
    ref = fd.HestonBatch(P, S, Vs, dts)
    ref.set_state(init)
    ref.configure("cs", 0.5)
    ref.step(6)
    tuned = fd.HestonBatch(P, S, Vs, dts)
    tuned.set_state(init)
    tuned.configure("cs", 0.5)
    before = tuned.get_packed()
    shapes = tuned.autotune(steps=1)
    assert set(shapes) == {0, 1}
    npt.assert_array_equal(tuned.get_packed(), before)
    assert tuned.get_shape(0) == shapes[0] and tuned.get_shape(1)[0] == shapes[1][0]
    tuned.step(6)
    scale = np.max(np.abs(ref.get_packed()))
    assert np.max(np.abs(tuned.get_packed() - ref.get_packed())) <= 1e-12 * scale


def test_config2_batch_subset_matches_oracle(fd):
    """The 4096-option Craig-Sneyd batch (uneven grids nS in 64..512, nv in
    32..256, nt = 200) in one engine; two options of it (the largest and a
    random one) against the oracle.

    These meshes put the first variance node ~1e-10 above v = 0, so the
    discrete operators carry entries ~1e10 next to values ~1e2 and results
    depend on the order of operations well above 1e-10: the oracle moves by
    ~1e-8 on the grid (and ~1e-12 relative at the price node) when only the
    order of its own explicit sum, of its band products, or of its
    tridiagonal elimination changes.  The grid bar is therefore
    max(1e-10, 4 x the largest of those three moves); the price node is
    held to the same absolute bar (its relative error is printed and
    recorded)."""
    from fd_adi_amd.workloads import batch_draws
    opts = batch_draws(4096, 2024)
    hb = fd.HestonBatch([o.params for o in opts], [o.spots for o in opts],
                        [o.vars for o in opts], [o.dt for o in opts])
    hb.set_state([o.payoff() for o in opts])
    hb.configure("cs", 0.5)
    hb.step(opts[0].nt)
    rng = np.random.default_rng(3)
    big = int(np.argmax([o.spots.size * o.vars.size for o in opts]))
    for k in [big, int(rng.integers(len(opts)))]:
        o = opts[k]
        hp = HestonParams(spot=o.S0, strike=o.K, r=o.r, v0=o.v0, T=o.T, kappa=o.kappa,
                          theta=o.theta, sigma=o.sigma, rho=o.rho)
        U0 = o.payoff()

        def solve(**kw):
            return HestonOracle(hp, o.spots, o.vars, **kw).solve_craigsneyd(o.nt, o.dt, U0, 0.5)
        want = solve()
        floor = max(np.max(np.abs(want - solve(firsts_reversed=True))),
                    np.max(np.abs(want - solve(apply_order="reversed"))),
                    np.max(np.abs(want - solve(tri_solver="thomas"))))
        tol = max(1e-10, FLOOR_FACTOR * floor)
        got = hb.state_view(k).to_numpy().reshape(o.spots.size, o.vars.size)
        err = float(np.max(np.abs(got - want)))
        dp = abs(got[o.idx] - want[o.idx])
        print("option %d (%dx%d): grid err %.3e tol %.3e floor %.3e; price abs %.3e rel %.3e"
              % (k, o.spots.size, o.vars.size, err, tol, floor, dp, dp / abs(want[o.idx])))
        record(err=err, floor=float(floor), tol=tol, ratio=err / floor, price_rel=dp / abs(want[o.idx]),
               case="config 2 option %d (%dx%d)" % (k, o.spots.size, o.vars.size))
        assert err <= tol and dp <= tol
