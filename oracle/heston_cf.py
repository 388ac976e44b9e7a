"""Semi-analytic Heston European call -- independent price check.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference checks its schemes against a COS-method pricer
(FiniteDifference/heston.py:490-693, rtol 1e-3 in
FiniteDifference/tests/test_HestonOptionGPU.py:74-112).  This module does not
restate that code; it prices with the Gil-Pelaez inversion of the Heston
characteristic function in the "little trap" form of Albrecher et al. (2007),
integrated with adaptive quadrature, which is an independent method that
fixes the same number.

    C = S P1 - K e^{-rT} P2,
    Pj = 1/2 + 1/pi * int_0^inf Re[ e^{-iu ln K} phi_j(u) / (iu) ] du
"""

import numpy as np
import scipy.integrate


def _phi(u, S, v0, r, T, kappa, theta, sigma, rho, j):
    x = np.log(S)
    if j == 1:
        uj, bj = 0.5, kappa - rho * sigma
    else:
        uj, bj = -0.5, kappa
    a = kappa * theta
    iu = 1j * u
    d = np.sqrt((rho * sigma * iu - bj) ** 2 - sigma ** 2 * (2 * uj * iu - u ** 2))
    g = (bj - rho * sigma * iu - d) / (bj - rho * sigma * iu + d)
    e = np.exp(-d * T)
    C = r * iu * T + a / sigma ** 2 * (
        (bj - rho * sigma * iu - d) * T - 2 * np.log((1 - g * e) / (1 - g)))
    D = (bj - rho * sigma * iu - d) / sigma ** 2 * ((1 - e) / (1 - g * e))
    return np.exp(C + D * v0 + iu * x)


def heston_call(S, K, r, v0, T, kappa, theta, sigma, rho):
    """European call price under Heston dynamics."""
    lnK = np.log(K)

    def integrand(u, j):
        return (np.exp(-1j * u * lnK) *
                _phi(u, S, v0, r, T, kappa, theta, sigma, rho, j) / (1j * u)).real

    P = []
    for j in (1, 2):
        # to infinity: for small (v0 + kappa theta T) / sigma the integrand
        # still matters beyond u = 200 (1e-5 of the price at the box's edge)
        val, _ = scipy.integrate.quad(integrand, 1e-10, np.inf, args=(j,),
                                      limit=4000, epsabs=1e-13, epsrel=1e-11)
        P.append(0.5 + val / np.pi)
    return S * P[0] - K * np.exp(-r * T) * P[1]


def bs_call(S, K, r, vol, T):
    """Black-Scholes closed form (FiniteDifference/blackscholes.py:42-58)."""
    from scipy.stats import norm
    d1 = (np.log(S / K) + (r + 0.5 * vol * vol) * T) / (vol * np.sqrt(T))
    d2 = d1 - vol * np.sqrt(T)
    return S * norm.cdf(d1) - K * np.exp(-r * T) * norm.cdf(d2)
