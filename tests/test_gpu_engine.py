"""GPU schemes vs the oracle on identical grids and parameters.

Bar (tests/parity.py): relative price error at the (S0, v0) node <= 1e-12
and max |V_gpu - V_oracle| over the grid <= max(1e-10, 4 x the oracle's own
rounding-order noise), for every scheme on both the fused and the
operator-algebra path; the reference's own GPU-vs-CPU bar was decimal 6-7
(FiniteDifference/tests/test_HestonOptionGPU.py:73,100).
"""

import numpy as np
import numpy.testing as npt
import pytest

pytestmark = pytest.mark.gpu

from oracle.engine import HestonOracle, HestonParams
from oracle.heston_cf import heston_call
from tests.parity import check as parity_check, alternates

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
    """Oracle result and the same solve under each rounding-order probe."""
    return fn(O), [fn(a) for a in alternates(O)]


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


@pytest.mark.parametrize("scheme", ["implicit", "douglas", "hundsdorferverwer", "smooth"])
def test_reference_case_150x80(fd, scheme):
    """The case of test_HestonOptionGPU.py:62-112: 150 x 80 grid, rho = 0.4,
    dt = 1/600 for the implicit scheme and 1/150 otherwise; each within the
    parity bar of the oracle and within rtol 1e-3 of the semi-analytic price
    (the reference's own tolerance, :74,87,99,111)."""
    G, O = make(150, 80, 0.4)
    dt = 1 / 600.0 if scheme == "implicit" else 1 / 150.0
    n = int(round(1.0 / dt))
    U0 = O.payoff()
    V = getattr(G, "solve_" + scheme)(n, dt, U0)
    W = oracle_both(O, lambda E: getattr(E, "solve_" + scheme)(n, dt, U0))
    check(V, W, O.idx)
    npt.assert_allclose(V[O.idx], heston_call(100, 100, 0.06, 0.04, 1.0, 1.0, 0.12, 0.3, 0.4),
                        rtol=1e-3)
    npt.assert_allclose(G.price, V[G.idx])


def test_dummy_step_operators_and_state(fd):
    """One Douglas step from arange(N) through the operator path
    (FiniteDifferenceEngineGPU.pyx:283-340) -- the reference's test_firsts,
    test_Les, test_Lis, test_Orig, test_Y and test_V
    (tests/test_FiniteDifferenceEngineGPU.py:288-356): every operator's bands
    read back from the device against the oracle's, and the state after the
    step against the oracle's Douglas step."""
    G, O = make(40, 24, 0.4)
    G.fused = False
    G.make_operator_templates()
    G.set_zero_derivative()
    G.scale_and_combine_operators()
    Firsts, Les, Lis, Orig, Y, V = G._dummy()
    dt, theta = 0.01, 0.5
    keys = list(G.operators)
    dims = sorted(k for k in keys if not isinstance(k, tuple))

    def same(dev, ref):
        host = dev.immigrate()
        if ref.is_foldable():
            ref.diagonalize()
        npt.assert_array_equal(host.offsets, ref.offsets)
        npt.assert_allclose(host.data, ref.data, rtol=1e-13,
                            atol=1e-15 * np.max(np.abs(ref.data)))
        if host.R is not None and ref.R is not None:
            npt.assert_allclose(host.R, ref.R, rtol=1e-13, atol=1e-15 * np.max(np.abs(ref.R)))
        else:
            assert not np.any(host.R) and not np.any(ref.R)

    for k, dev in zip(keys, Firsts):
        same(dev, O.operators[k] * dt)
    Les_o, Lis_o = O._les_lis(dt, theta, fold=False)
    for dev, ref in zip(Les, Les_o):
        same(dev, ref)
    for dev, ref in zip(Lis, Lis_o):
        same(dev, ref)
    start = np.arange(G.shape[0] * G.shape[1], dtype=float).reshape(G.shape)
    npt.assert_array_equal(Orig, start)
    npt.assert_array_equal(Y, start)
    check(V, oracle_both(O, lambda E: E.solve_douglas(1, dt, start, theta=theta)), O.idx)


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


@pytest.mark.parametrize("scheme", ["douglas", "hundsdorferverwer"])
def test_unsnapped_payoff_from_engine_domain(fd, scheme):
    """S0 != K (the command line's default strike 99): the payoff is built on
    the spot mesh as generated, then S0 is snapped onto its nearest node in
    the array shared with grid.mesh[0] (FiniteDifferenceEngineGPU.pyx:682-686).
    The solve starts from the engine's own domain (no U0 supplied) and the
    oracle starts from the unsnapped payoff on the snapped mesh."""
    from fd_adi_amd.FiniteDifferenceEngineGPU import HestonFiniteDifferenceEngine
    from fd_adi_amd.Grid import Grid
    from fd_adi_amd.Option import HestonOption
    from fd_adi_amd.utils import sinh_space, exponential_space
    K, S0, nS, nv = 99.0, 100.0, 100, 60
    kw = dict(spot=S0, strike=K, interest_rate=0.06, variance=0.04, tenor=1.0,
              mean_reversion=1.0, mean_variance=0.04, vol_of_variance=0.4, correlation=-0.3)
    raw = sinh_space(K, 1500.0, 7.0, nS)
    vars_ = exponential_space(0.0, 0.04, 10.0, 4.0, nv)
    G = HestonFiniteDifferenceEngine(HestonOption(**kw), nspots=nS, nvols=nv)
    i = G.idx[0]
    assert G.grid.mesh[0] is G.spots and G.spots[i] == S0 and raw[i] != S0
    npt.assert_array_equal(np.delete(G.spots, i), np.delete(raw, i))
    npt.assert_array_equal(G.vars, vars_)
    U0 = np.maximum(raw - K, 0.0)[:, None] + 0.0 * vars_[None, :]
    npt.assert_array_equal(G.grid.domain[-1], U0)
    assert U0[G.idx] != max(S0 - K, 0.0)
    p = HestonParams(spot=S0, strike=K, r=0.06, v0=0.04, T=1.0, kappa=1.0, theta=0.04,
                     sigma=0.4, rho=-0.3)
    O = HestonOracle(p, G.spots, G.vars)
    V = getattr(G, "solve_" + scheme)(50, 0.02)
    W = oracle_both(O, lambda E: getattr(E, "solve_" + scheme)(50, 0.02, U0))
    check(V, W, O.idx)
    assert G.idx == O.idx and G.price == V[G.idx]

    # grid= branch: the caller's Grid keeps its payoff and its mesh is snapped
    g = Grid([raw.copy(), vars_.copy()], initializer=lambda *x: np.maximum(x[0] - K, 0))
    G2 = HestonFiniteDifferenceEngine(HestonOption(**kw), grid=g)
    assert g.mesh[0][i] == S0 and G2.spots is g.mesh[0]
    npt.assert_array_equal(g.domain[-1], U0)
    V2 = getattr(G2, "solve_" + scheme)(50, 0.02)
    npt.assert_array_equal(V2, V)


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "operators"])
def test_device_scalar_dt_theta(fd, fused):
    """The device variants take dt (and theta for HV) as device scalars, as
    the reference's do (FiniteDifferenceEngineGPU.pyx:379,502,582):
    solve_implicit_, solve_hundsdorferverwer_ and solve_smooth_ with
    SizedArrayPtr dt / theta against the oracle; then a host-scalar solve on
    the same engine (the configuration switches back to host values)."""
    from fd_adi_amd.SizedArrayPtr import SizedArrayPtr
    G, O = make(40, 20, -0.4)
    G.fused = fused
    U0 = O.payoff()
    dt, th = SizedArrayPtr(np.array([0.05])), SizedArrayPtr(np.array([0.5]))
    V = SizedArrayPtr(U0)
    G.solve_hundsdorferverwer_(10, dt, V, th)
    check(V.to_numpy(), oracle_both(O, lambda E: E.solve_hundsdorferverwer(10, 0.05, U0)), O.idx)
    V = SizedArrayPtr(U0)
    G.solve_implicit_(8, dt, V)
    check(V.to_numpy(), oracle_both(O, lambda E: E.solve_implicit(8, 0.05, U0)), O.idx)
    V = SizedArrayPtr(U0)
    G.solve_smooth_(12, dt, V)
    check(V.to_numpy(), oracle_both(O, lambda E: E.solve_smooth(12, 0.05, U0)), O.idx)
    V = G.solve_hundsdorferverwer(10, 0.04, U0, theta=0.6)
    check(V, oracle_both(O, lambda E: E.solve_hundsdorferverwer(10, 0.04, U0, theta=0.6)),
          O.idx)
