"""Pins of the oracle's schemes (CPU only): an independent semi-analytic
price (the reference checks its schemes the same way at rtol 1e-3,
FiniteDifference/tests/test_HestonOptionGPU.py:74-112), the Black-Scholes
limit, observed time-convergence orders and the Dirichlet invariant."""

import numpy as np
import numpy.testing as npt
import pytest

from oracle.engine import HestonOracle, HestonParams
from oracle.heston_cf import heston_call, bs_call
from fd_adi_amd.utils import sinh_space, exponential_space

SCHEMES = ["implicit", "douglas", "hundsdorferverwer", "craigsneyd", "craigsneyd2"]


def reference_engine(nS=150, nv=80, rho=0.4):
    """The reference test's Heston case (test_HestonOptionGPU.py:29-48)."""
    p = HestonParams(spot=100, strike=100, r=0.06, v0=0.04, T=1.0, kappa=1.0, theta=0.12,
                     sigma=0.3, rho=rho)
    spots = sinh_space(100, 1500, 7.0, nS)
    spots[np.argmin(np.abs(spots - 100))] = 100
    vars_ = exponential_space(0, 0.04, 10.0, 4.0, nv)
    return HestonOracle(p, spots, vars_)


def test_cf_pricer_black_scholes_limit():
    # sigma -> 0 with v0 = theta: constant variance -> Black-Scholes at vol sqrt(v0)
    for K in (80.0, 100.0, 120.0):
        h = heston_call(100, K, 0.05, 0.04, 1.0, 2.0, 0.04, 1e-4, 0.0)
        npt.assert_allclose(h, bs_call(100, K, 0.05, 0.2, 1.0), rtol=1e-6)


def test_cf_pricer_put_call_parity_and_bounds():
    S, K, r, T = 100.0, 110.0, 0.03, 0.7
    c = heston_call(S, K, r, 0.05, T, 1.5, 0.06, 0.5, -0.6)
    assert max(S - K * np.exp(-r * T), 0) < c < S
    # the call is increasing in v0 and decreasing in K
    assert heston_call(S, K, r, 0.08, T, 1.5, 0.06, 0.5, -0.6) > c
    assert heston_call(S, K + 5, r, 0.05, T, 1.5, 0.06, 0.5, -0.6) < c


@pytest.mark.parametrize("scheme", SCHEMES)
def test_schemes_vs_semi_analytic(scheme):
    E = reference_engine()
    ans = heston_call(100, 100, 0.06, 0.04, 1.0, 1.0, 0.12, 0.3, 0.4)
    dt = 1 / 600.0 if scheme == "implicit" else 1 / 150.0
    V = getattr(E, "solve_" + scheme)(round(1 / dt), dt, E.payoff())
    npt.assert_allclose(V[E.idx], ans, rtol=1e-3)
    assert np.all(V[0, :] == 0.0)


def test_smooth_vs_semi_analytic():
    E = reference_engine()
    ans = heston_call(100, 100, 0.06, 0.04, 1.0, 1.0, 0.12, 0.3, 0.4)
    V = E.solve_smooth(150, 1 / 150.0, E.payoff())
    npt.assert_allclose(V[E.idx], ans, rtol=1e-3)


def test_black_scholes_limit_of_the_pde():
    p = HestonParams(spot=100, strike=100, r=0.05, v0=0.04, T=1.0, kappa=2.0, theta=0.04,
                     sigma=1e-3, rho=0.0)
    spots = sinh_space(100, 800, 7.0, 160)
    spots[np.argmin(np.abs(spots - 100))] = 100
    vars_ = exponential_space(0, 0.04, 0.4, 2.0, 24)
    E = HestonOracle(p, spots, vars_)
    V = E.solve_hundsdorferverwer(100, 0.01, E.payoff())
    npt.assert_allclose(V[E.idx], bs_call(100, 100, 0.05, 0.2, 1.0), rtol=1e-3)


@pytest.mark.parametrize("scheme,order", [("implicit", 1), ("hundsdorferverwer", 2),
                                          ("craigsneyd", 2), ("craigsneyd2", 2)])
def test_time_convergence_order(scheme, order):
    E = reference_engine(nS=40, nv=20, rho=-0.5)
    U0 = E.payoff()
    vals = []
    for n in (100, 200, 400):
        vals.append(getattr(E, "solve_" + scheme)(n, 1.0 / n, U0)[E.idx])
    e1, e2 = abs(vals[0] - vals[1]), abs(vals[1] - vals[2])
    rate = np.log2(e1 / e2)
    assert abs(rate - order) < 0.2, (vals, rate)


def test_float32_rate_quirk_is_small_but_visible():
    """The reference CPU engine rounds -r/2 to single precision
    (BandedOperator.pyx:713, `float other`); the GPU engine does not."""
    E = reference_engine(nS=40, nv=20)
    F = HestonOracle(E.p, E.spots, E.vars, r_split_float32=True)
    U0 = E.payoff()
    a = E.solve_douglas(20, 0.05, U0)[E.idx]
    b = F.solve_douglas(20, 0.05, U0)[E.idx]
    assert a != b and abs(a - b) / abs(a) < 1e-8


def test_rounding_order_probes_match_default():
    """The oracle's rounding-order probes (reversed explicit sum, reversed
    band products, Thomas elimination instead of LAPACK) compute the same
    scheme: they agree with the default oracle to rounding on a small grid,
    and the Thomas solve agrees with a dense solve."""
    from fd_adi_amd.utils import sinh_space, exponential_space
    from oracle.engine import HestonOracle, HestonParams
    sp = sinh_space(100, 500, 7.0, 40)
    vv = exponential_space(0, 0.04, 0.2, 4.0, 20)
    hp = HestonParams(spot=100, strike=100, r=0.06, v0=0.04, T=1.0, kappa=1.0, theta=0.04,
                      sigma=0.3, rho=-0.5)
    base = HestonOracle(hp, sp, vv)
    U0 = base.payoff()
    W = base.solve_craigsneyd(10, 0.05, U0, 0.5)
    for kw in (dict(firsts_reversed=True), dict(apply_order="reversed"),
               dict(tri_solver="thomas")):
        Wp = HestonOracle(hp, sp, vv, **kw).solve_craigsneyd(10, 0.05, U0, 0.5)
        d = np.max(np.abs(Wp - W)) / np.max(np.abs(W))
        assert 0 < d < 1e-12, (kw, d)
    # the Thomas path on one folded variance operator against a dense solve
    E = HestonOracle(hp, sp, vv, tri_solver="thomas")
    Les, Lis = E._les_lis(0.05, 0.5, fold=True)
    Li = Lis[1]
    rhs = np.random.default_rng(0).normal(size=U0.shape)
    x_thomas = Li.solve(rhs)
    Li.solver = "lapack"
    x_lapack = Li.solve(rhs)
    npt.assert_allclose(x_thomas, x_lapack, rtol=0, atol=1e-12 * np.max(np.abs(x_lapack)))
