"""The committed golden fixtures (CPU only).

* the oracle reproduces every fixture's value grid on its stored inputs
  (a regression pin of the restatement);
* the host mesh generators reproduce the stored meshes;
* every fixture's price agrees with the independent semi-analytic Heston
  price (rtol 5e-4 on the 150 x 150 headline case, 1e-2 on the 40 x 24 ones,
  3e-2 for the first-order implicit scheme;
  the reference checks its own schemes this way at rtol 1e-3,
  FiniteDifference/tests/test_HestonOptionGPU.py:74-112);
* the up-and-out barrier price lies below the vanilla price on the same
  parameters and its barrier row is zero to rounding.
"""

import numpy as np
import numpy.testing as npt
import pytest

from oracle.engine import HestonOracle, HestonParams
from fd_adi_amd.utils import sinh_space, exponential_space
from tests.golden_cases import names, load, solve

NAMES = names()


def params(c):
    p = c.p
    return HestonParams(spot=p["spot"], strike=p["strike"], r=p["r"], v0=p["v0"], T=p["T"],
                        kappa=p["kappa"], theta=p["theta"], sigma=p["sigma"], rho=p["rho"],
                        barrier=c.barrier)


def test_fixtures_present():
    assert len(NAMES) == 8, NAMES


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    c = load(name)
    O = HestonOracle(params(c), c.spots, c.vars)
    npt.assert_array_equal(O.payoff(), c.U0)
    assert O.idx == c.idx
    V = solve(O, c, c.U0)
    npt.assert_allclose(V, c.V, rtol=0, atol=1e-11)


@pytest.mark.parametrize("name", [n for n in NAMES if not n.startswith("config0")])
def test_oracle_reproduces_floors(name):
    """The stored rounding-order floors come back from the probes."""
    from tests.parity import PROBES
    c = load(name)
    V = solve(HestonOracle(params(c), c.spots, c.vars), c, c.U0)
    got = [np.max(np.abs(V - solve(HestonOracle(params(c), c.spots, c.vars, **kw), c, c.U0)))
           for kw in PROBES]
    npt.assert_allclose(got, c.floors, rtol=1e-6, atol=1e-14)
    assert np.max(c.floors) > 0


@pytest.mark.parametrize("name", NAMES)
def test_meshes_regenerate(name):
    c = load(name)
    p = c.p
    if c.barrier is None:
        spots = sinh_space(p["strike"], c.meta["spot_max"], 7.0, len(c.spots))
    else:
        spots = exponential_space(0.0, p["strike"], c.barrier, 1.0 / 3, len(c.spots))
    spots[np.argmin(np.abs(spots - p["spot"]))] = p["spot"]
    vars_ = exponential_space(0.0, p["v0"], c.meta["var_max"], 4.0, len(c.vars))
    npt.assert_allclose(spots, c.spots, rtol=1e-13, atol=1e-12)
    npt.assert_allclose(vars_, c.vars, rtol=1e-13, atol=1e-15)
    assert c.spots[c.idx[0]] == p["spot"] and abs(c.vars[c.idx[1]] - p["v0"]) < 1e-12


@pytest.mark.parametrize("name", [n for n in NAMES if "barrier" not in n])
def test_fixture_price_vs_semi_analytic(name):
    c = load(name)
    # fully implicit is first order in time: 12 steps leave ~2%
    rtol = 5e-4 if name.startswith("config0") else 3e-2 if name.startswith("implicit") else 1e-2
    npt.assert_allclose(c.V[c.idx], c.price_cf, rtol=rtol)


def test_barrier_fixture_below_vanilla():
    c = load("hv_barrier_48x24")
    # the barrier row is Dirichlet zero in the spot sweeps; the last variance
    # sweep leaves rounding-level values there
    assert np.max(np.abs(c.V[-1, :])) < 1e-12
    vanilla = load("hv_40x24")
    assert 0.0 < c.V[c.idx] < vanilla.price_cf
