"""GPU schemes vs the oracle on identical grids and parameters.

Bar (tests/parity.py): relative price error at the (S0, v0) node <= 1e-12
and max |V_gpu - V_oracle| over the grid <= max(1e-10, 10 x the oracle's own
reordering noise), for every scheme on both the fused and the
operator-algebra path; the reference's own GPU-vs-CPU bar was decimal 6-7
(FiniteDifference/tests/test_HestonOptionGPU.py:73,100).
"""

import numpy as np
import numpy.testing as npt
import pytest

pytestmark = pytest.mark.gpu

from oracle.engine import HestonOracle, HestonParams
from oracle.heston_cf import heston_call
from tests.parity import check as parity_check, reordered

SCHEMES = ["implicit", "douglas", "hundsdorferverwer", "craigsneyd", "craigsneyd2"]


@pytest.fixture(scope="module")
def fd(gpu):
    import fd_adi_amd
    return fd_adi_amd


def make(nS, nv, rho, barrier=None, r=0.06, kappa=1.0, theta=0.12, sigma=0.3, v0=0.04,
         spot_max=1500.0, var_max=10.0):
    from fd_adi_amd.FiniteDifferenceEngineGPU import HestonFiniteDifferenceEngine
    from fd_adi_amd.Option import HestonOption, HestonBarrierOption
    kw = dict(spot=100, strike=100, interest_rate=r, variance=v0, tenor=1.0,
              mean_reversion=kappa, mean_variance=theta, vol_of_variance=sigma, correlation=rho)
    opt = HestonBarrierOption(top=(False, barrier), **kw) if barrier else HestonOption(**kw)
    G = HestonFiniteDifferenceEngine(opt, nspots=nS, nvols=nv, spot_max=spot_max,
                                     var_max=var_max)
    p = HestonParams(spot=100, strike=100, r=r, v0=v0, T=1.0, kappa=kappa, theta=theta,
                     sigma=sigma, rho=rho, barrier=barrier)
    O = HestonOracle(p, G.spots, G.vars)
    return G, O


def oracle_both(O, fn):
    """Oracle result and the same oracle with the explicit sum reordered."""
    return fn(O), fn(reordered(O))


def check(V, pair, idx):
    return parity_check(V, pair[0], pair[1], idx)


def run(E, scheme, n, dt, U0, theta=0.5):
    f = getattr(E, "solve_" + scheme)
    if scheme == "implicit":
        return f(n, dt, U0)
    return f(n, dt, U0, theta=theta)


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("rho", [0.0, -0.7])
def test_fused_scheme_matches_oracle(fd, scheme, rho):
    G, O = make(48, 24, rho)
    U0 = O.payoff()
    V = run(G, scheme, 20, 0.05, U0)
    W = oracle_both(O, lambda E: run(E, scheme, 20, 0.05, U0))
    print(scheme, rho, check(V, W, O.idx))


@pytest.mark.parametrize("scheme", SCHEMES)
def test_operator_path_matches_oracle(fd, scheme):
    G, O = make(40, 20, 0.4)
    G.fused = False
    U0 = O.payoff()
    V = run(G, scheme, 10, 0.05, U0)
    W = oracle_both(O, lambda E: run(E, scheme, 10, 0.05, U0))
    print(scheme, check(V, W, O.idx))


def test_reference_case_hv_and_price(fd):
    """test_HestonOptionGPU.py:92-100 case: 150 x 80, dt = 1/150, HV."""
    G, O = make(150, 80, 0.4)
    U0 = O.payoff()
    V = G.solve_hundsdorferverwer(150, 1 / 150.0, U0)
    W = oracle_both(O, lambda E: E.solve_hundsdorferverwer(150, 1 / 150.0, U0))
    check(V, W, O.idx)
    npt.assert_allclose(V[O.idx], heston_call(100, 100, 0.06, 0.04, 1.0, 1.0, 0.12, 0.3, 0.4),
                        rtol=1e-3)
    npt.assert_allclose(G.price, V[G.idx])


def test_smooth_matches_oracle(fd):
    G, O = make(64, 32, -0.3)
    U0 = O.payoff()
    V = G.solve_smooth(30, 1 / 30.0, U0)
    W = oracle_both(O, lambda E: E.solve_smooth(30, 1 / 30.0, U0))
    check(V, W, O.idx)


@pytest.mark.parametrize("scheme", ["douglas", "hundsdorferverwer", "craigsneyd"])
def test_barrier_matches_oracle(fd, scheme):
    G, O = make(60, 30, -0.5, barrier=170.0)
    U0 = O.payoff()
    V = run(G, scheme, 20, 0.05, U0)
    W = oracle_both(O, lambda E: run(E, scheme, 20, 0.05, U0))
    check(V, W, O.idx)
    # the barrier row is Dirichlet for the spot sweep; the final variance sweep
    # leaves roundoff there, as in the oracle
    assert np.max(np.abs(V[-1])) < 1e-12 and np.all(V[0] == 0)


def test_device_variant_in_place(fd):
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    G, O = make(40, 20, 0.2)
    U0 = O.payoff()
    V = SizedArrayPtr(U0)
    G.solve_craigsneyd_(10, 0.05, V, 0.5)
    check(V.to_numpy(), oracle_both(O, lambda E: E.solve_craigsneyd(10, 0.05, U0)), O.idx)


def test_config0_full_size(fd):
    """150 x 150, nt = 150: implicit(4, dt/2) then Douglas(148, dt, theta=0.5)."""
    from fd_adi_amd.utils import sinh_space, exponential_space
    G, O = make(150, 150, 0.0, r=0.06, kappa=1.0, theta=0.04, sigma=0.2, v0=0.04,
                spot_max=500.0, var_max=0.2)
    dt = 1.0 / 150
    U0 = O.payoff()
    V = G.solve_implicit(4, dt / 2, U0)
    V = G.solve_douglas(148, dt, V, theta=0.5)
    W = oracle_both(O, lambda E: E.solve_douglas(148, dt, E.solve_implicit(4, dt / 2, U0),
                                                 theta=0.5))
    check(V, W, O.idx)
    npt.assert_allclose(V[O.idx], heston_call(100, 100, 0.06, 0.04, 1.0, 1.0, 0.04, 0.2, 0.0),
                        rtol=2e-3)
