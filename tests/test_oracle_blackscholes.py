"""The one-dimensional Black-Scholes oracle on the reference's own case
(FiniteDifference/tests/test_BlackScholesOption.py:24-75: S0 = 100, K = 99,
r = 0.06, variance 0.04, T = 1, 150 sinh-spaced spots up to 5000 with
density 1, dt = 1/150): every scheme within rtol 1e-3 of the closed form, as
the reference asserts, and the combined operator tridiagonal."""

import numpy as np
import numpy.testing as npt
import pytest

from oracle.blackscholes import BlackScholesOracle, BlackScholesParams
from oracle.heston_cf import bs_call
from fd_adi_amd.utils import sinh_space


def reference_case(**kw):
    p = BlackScholesParams(spot=100.0, strike=99.0, r=0.06, variance=0.04, T=1.0)
    return BlackScholesOracle(p, sinh_space(100.0, 5000.0, 1.0, 150), **kw)


ANSWER = bs_call(100.0, 99.0, 0.06, 0.2, 1.0)


def test_operator_is_tridiagonal_with_the_boundaries():
    O = reference_case()
    assert list(O.operators) == [0]
    A = O.operators[0]
    assert A.is_tridiagonal()
    assert A.dirichlet[0] == 0.0 and A.dirichlet[1] is None
    # interior row: r S d/dS + 0.5 var S^2 d2/dS2 - r applied to S is r S - r S = 0
    S = O.spots
    y = A.apply(S)
    npt.assert_allclose(y[1:-1], 0.0, atol=1e-9 * S.max())


@pytest.mark.parametrize("scheme", ["implicit", "douglas", "smooth", "hundsdorferverwer",
                                    "craigsneyd"])
def test_schemes_match_closed_form(scheme):
    O = reference_case()
    dt = 1.0 / 150
    f = getattr(O, "solve_" + scheme)
    V = f(150, dt, O.payoff())
    npt.assert_allclose(V[O.idx], ANSWER, rtol=1e-3)
