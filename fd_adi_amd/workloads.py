"""Synthetic Heston workloads (parameter draws and meshes).

* ``single(...)``: one European call on a sinh (spot) x power-law (variance)
  grid, built exactly as the engine builds it
  (FiniteDifference/FiniteDifferenceEngineGPU.pyx:658-686).
* ``label_draws(n, seed)``: surrogate rows on the label meshes (below).
* ``batch_draws(n, seed, ...)``: the batched workload: per option
  kappa~U[0.5,5], theta~U[0.01,0.16], sigma~U[0.1,1], rho~U[-0.9,0],
  r~U[0,0.08], T~U[0.1,3], K/S0~U[0.7,1.3] (S0=100), v0=theta (or drawn),
  nS~U{64..512}, nv~U{32..256}; spot mesh sinh_space(K, 5K, 7, nS) with S0
  snapped onto a node, variance mesh exponential_space(0, v0, 5*max(theta,
  v0), 4, nv).
"""

import numpy as np

from .utils import sinh_space, exponential_space, nearest_index


class Option1(object):
    """Plain container: parameters, meshes, step and the price node."""

    def __init__(self, S0, K, r, v0, T, kappa, theta, sigma, rho, spots, vars, nt):
        self.S0, self.K, self.r, self.v0, self.T = S0, K, r, v0, T
        self.kappa, self.theta, self.sigma, self.rho = kappa, theta, sigma, rho
        self.spots, self.vars, self.nt = spots, vars, nt
        self.dt = T / nt
        self.idx = (nearest_index(spots, S0), nearest_index(vars, v0))

    @property
    def params(self):
        return dict(r=self.r, kappa=self.kappa, theta=self.theta, sigma=self.sigma,
                    rho=self.rho)

    def payoff(self):
        return np.maximum(self.spots[:, None] - self.K, 0.0) + 0.0 * self.vars[None, :]


def single(nS, nv, nt, rho, S0=100.0, K=100.0, r=0.06, v0=0.04, T=1.0, kappa=1.0,
           theta=0.04, sigma=0.2, spot_max=None, var_max=None):
    spot_max = 5 * K if spot_max is None else spot_max
    var_max = 5 * theta if var_max is None else var_max
    spots = sinh_space(K, spot_max, 7.0, nS)
    spots[np.argmin(np.abs(spots - S0))] = S0
    vars_ = exponential_space(0.0, v0, var_max, 4.0, nv)
    return Option1(S0, K, r, v0, T, kappa, theta, sigma, rho, spots, vars_, nt)


def batch_draws(n, seed, nS_range=(64, 512), nv_range=(32, 256), nt=200, draw_v0=False,
                box_scale=0.0):
    """``n`` options drawn from ``seed``.  ``box_scale`` > 0 widens every
    parameter range by that fraction of its width on both sides."""
    rng = np.random.default_rng(seed)

    def U(lo, hi, size=None):
        w = hi - lo
        return rng.uniform(lo - box_scale * w, hi + box_scale * w, size)

    out = []
    for _ in range(n):
        kappa = max(U(0.5, 5.0), 0.05)
        theta = max(U(0.01, 0.16), 0.002)
        sigma = max(U(0.1, 1.0), 0.02)
        rho = min(max(U(-0.9, 0.0), -0.99), 0.5)
        r = max(U(0.0, 0.08), 0.0)
        T = max(U(0.1, 3.0), 0.05)
        K = 100.0 * max(U(0.7, 1.3), 0.3)
        v0 = max(U(0.01, 0.16), 0.002) if draw_v0 else theta
        nS = int(rng.integers(nS_range[0], nS_range[1] + 1))
        nv = int(rng.integers(nv_range[0], nv_range[1] + 1))
        spots = sinh_space(K, 5 * K, 7.0, nS)
        spots[np.argmin(np.abs(spots - 100.0))] = 100.0
        vars_ = exponential_space(0.0, v0, 5 * max(theta, v0), 4.0, nv)
        out.append(Option1(100.0, K, r, v0, T, kappa, theta, sigma, rho, spots, vars_, nt))
    return out


def label_draws(n, seed, nS=256, nv=128, nt=200):
    """``n`` surrogate parameter rows drawn from ``seed`` as options on the
    label meshes (spot mesh K x one unit sinh mesh, fixed power-law variance
    mesh; S0 = 100 and v0 off the nodes): the options the surrogate's labels
    and its ADI fallback solve."""
    from .surrogate import draw_params, label_grids
    P = draw_params(n, seed)
    S, V = label_grids(P, nS, nv)
    return [Option1(100.0, 100.0 * p[7], p[5], p[4], p[6], p[0], p[1], p[2], p[3], s, v, nt)
            for p, s, v in zip(P, S, V)]
