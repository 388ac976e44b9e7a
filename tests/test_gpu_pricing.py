"""Device semi-analytic Heston call (fdadi_heston_cf_call) against the
oracle's independent Gil-Pelaez pricer (two integrals, adaptive quadrature
to infinity) and against properties the price must have.

The reference checks its schemes with a semi-analytic COS pricer
(FiniteDifference/heston.py:490-693); this one is Lewis's single integral with
fixed Gauss-Legendre quadrature, so agreement pins both.
"""

import numpy as np
import numpy.testing as npt
import pytest

from oracle.heston_cf import heston_call as oracle_call, bs_call

pytestmark = pytest.mark.gpu


def draws(n, seed):
    rng = np.random.default_rng(seed)
    lo = np.array([0.5, 0.01, 0.1, -0.9, 0.01, 0.0, 0.1, 0.7])
    hi = np.array([5.0, 0.16, 1.0, 0.0, 0.16, 0.08, 3.0, 1.3])
    return rng.uniform(lo, hi, (n, 8))   # kappa theta sigma rho v0 r T K/S0


def test_matches_oracle_over_the_parameter_box(gpu):
    from fd_adi_amd.pricing import heston_call
    P = draws(120, 3)
    kappa, theta, sigma, rho, v0, r, T, m = P.T
    got = heston_call(100.0, 100.0 * m, r, v0, T, kappa, theta, sigma, rho)
    ref = np.array([oracle_call(100.0, 100.0 * p[7], p[5], p[4], p[6], p[0], p[1], p[2], p[3])
                    for p in P])
    err = np.abs(got - ref)
    print("max abs err %.3e" % err.max())
    assert err.max() < 1e-8


def test_reference_case_and_positive_correlation(gpu):
    from fd_adi_amd.pricing import heston_call
    # the reference test's Heston case (test_HestonOptionGPU.py:29-48) and rho > 0
    for rho in (0.4, -0.4, 0.0):
        got = heston_call(100.0, 100.0, 0.06, 0.04, 1.0, 1.0, 0.12, 0.3, rho)
        ref = oracle_call(100.0, 100.0, 0.06, 0.04, 1.0, 1.0, 0.12, 0.3, rho)
        npt.assert_allclose(got, ref, rtol=0, atol=1e-9)


def test_black_scholes_limit(gpu):
    from fd_adi_amd.pricing import heston_call
    # constant variance (v0 = theta, tiny sigma): Black-Scholes at vol sqrt(v0);
    # the Heston correction is O(sigma^2) ~ 1e-6 at sigma = 1e-3
    for K in (80.0, 100.0, 120.0):
        got = heston_call(100.0, K, 0.05, 0.04, 1.0, 2.0, 0.04, 1e-3, 0.0)
        npt.assert_allclose(got, bs_call(100.0, K, 0.05, 0.2, 1.0), rtol=0, atol=1e-5)


def test_homogeneity_bounds_and_monotonicity(gpu):
    from fd_adi_amd.pricing import heston_call
    K = np.linspace(60.0, 140.0, 41)
    c = heston_call(100.0, K, 0.03, 0.05, 2.0, 1.5, 0.08, 0.6, -0.7)
    c2 = heston_call(250.0, 2.5 * K, 0.03, 0.05, 2.0, 1.5, 0.08, 0.6, -0.7)
    npt.assert_allclose(c2, 2.5 * c, rtol=1e-11, atol=1e-9)
    assert np.all(np.diff(c) < 0)                        # decreasing in K
    assert np.all(np.diff(c, 2) > -1e-9)                 # convex in K
    lower = np.maximum(100.0 - K * np.exp(-0.03 * 2.0), 0.0)
    assert np.all(c >= lower - 1e-9) and np.all(c <= 100.0)


def test_invalid_rows_give_nan_and_empty_input(gpu):
    from fd_adi_amd.pricing import heston_call
    got = heston_call(100.0, 100.0, 0.05, 0.04, np.array([1.0, 0.0, -1.0]), 2.0, 0.04, 0.3, 0.0)
    assert np.isfinite(got[0]) and np.isnan(got[1]) and np.isnan(got[2])
    assert heston_call(np.zeros(0), 100.0, 0.05, 0.04, 1.0, 2.0, 0.04, 0.3, 0.0).shape == (0,)
