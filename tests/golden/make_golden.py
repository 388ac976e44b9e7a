"""Write the golden fixtures of the Heston schemes (tests/golden/*.npz).

Each fixture holds the inputs of one solve (parameters, meshes, payoff, steps,
time step, theta, scheme) and the oracle's value grid for it, together with
``floors``: how far the oracle's grid moves (max abs) when only the order of
its floating-point operations changes -- explicit operators summed in the
opposite order (the reference sums them in Python 2 dict order), band
products summed in reverse, Thomas elimination instead of LAPACK -- which is
the rounding floor a parity test must allow (tests/parity.py), and the
semi-analytic Heston call price at (S0, v0).

The oracle is the NumPy/SciPy restatement of the reference CPU engine
(FiniteDifference/FiniteDifferenceEngine.pyx:616-1016); the meshes come from
the restated generators (FiniteDifference/utils.py:102-190) the way the GPU
engine builds them (FiniteDifference/FiniteDifferenceEngineGPU.pyx:658-686).

    python tests/golden/make_golden.py        # rewrites every fixture
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.engine import HestonOracle, HestonParams  # noqa: E402
from oracle.heston_cf import heston_call  # noqa: E402
from tests.parity import PROBES  # noqa: E402
from fd_adi_amd.utils import sinh_space, exponential_space  # noqa: E402

# name: (nS, nv, scheme, n, dt, theta, parameters, spot_max, var_max, barrier)
BASE = dict(spot=100.0, strike=100.0, r=0.06, v0=0.04, T=1.0, kappa=1.0, theta=0.12,
            sigma=0.3, rho=-0.5)
CASES = {
    "implicit_40x24": (40, 24, "implicit", 12, 1 / 12.0, 1.0, {}, 1500.0, 10.0, None),
    "douglas_40x24": (40, 24, "douglas", 12, 1 / 12.0, 0.5, {}, 1500.0, 10.0, None),
    "hv_40x24": (40, 24, "hundsdorferverwer", 12, 1 / 12.0, 0.5, {}, 1500.0, 10.0, None),
    "cs_40x24": (40, 24, "craigsneyd", 12, 1 / 12.0, 0.5, {}, 1500.0, 10.0, None),
    "mcs_40x24": (40, 24, "craigsneyd2", 12, 1 / 12.0, 1 / 3.0, {}, 1500.0, 10.0, None),
    "smooth_40x24": (40, 24, "smooth", 12, 1 / 12.0, 0.6, {}, 1500.0, 10.0, None),
    "hv_barrier_48x24": (48, 24, "hundsdorferverwer", 12, 1 / 12.0, 0.5, {}, 1500.0, 10.0,
                         170.0),
    # the reference's headline single-option case: Douglas theta 0.5 after two
    # smoothing steps (four implicit half steps), 150 x 150, nt 150, rho 0
    "config0_150x150": (150, 150, "douglas_smoothed", 150, 1 / 150.0, 0.5,
                        dict(theta=0.04, sigma=0.2, rho=0.0, v0=0.04), 500.0, 0.2, None),
}


def meshes(p, nS, nv, spot_max, var_max, barrier):
    if barrier is None:
        spots = sinh_space(p["strike"], spot_max, 7.0, nS)
    else:
        # up-and-out: cubic mesh ending on the barrier
        # (FiniteDifference/FiniteDifferenceEngineGPU.pyx:670-678)
        spots = exponential_space(0.0, p["strike"], barrier, 1.0 / 3, nS)
    spots[np.argmin(np.abs(spots - p["spot"]))] = p["spot"]
    vars_ = exponential_space(0.0, p["v0"], var_max, 4.0, nv)
    return spots, vars_


def solve(O, scheme, n, dt, theta, U0):
    if scheme == "implicit":
        return O.solve_implicit(n, dt, U0)
    if scheme == "smooth":
        return O.solve_smooth(n, dt, U0)
    if scheme == "douglas_smoothed":
        V = O.solve_implicit(4, dt * 0.5, U0)
        return O.solve_douglas(n - 2, dt, V, theta=theta)
    return getattr(O, "solve_" + scheme)(n, dt, U0, theta=theta)


def make(name):
    nS, nv, scheme, n, dt, theta, over, spot_max, var_max, barrier = CASES[name]
    p = dict(BASE, **over)
    spots, vars_ = meshes(p, nS, nv, spot_max, var_max, barrier)
    hp = HestonParams(spot=p["spot"], strike=p["strike"], r=p["r"], v0=p["v0"], T=p["T"],
                      kappa=p["kappa"], theta=p["theta"], sigma=p["sigma"], rho=p["rho"],
                      barrier=barrier)
    O = HestonOracle(hp, spots, vars_)
    U0 = O.payoff()
    V = solve(O, scheme, n, dt, theta, U0)
    floors = np.array([np.max(np.abs(V - solve(HestonOracle(hp, spots, vars_, **kw), scheme, n,
                                                 dt, theta, U0))) for kw in PROBES])
    cf = np.nan if barrier is not None else heston_call(
        p["spot"], p["strike"], p["r"], p["v0"], p["T"], p["kappa"], p["theta"], p["sigma"],
        p["rho"])
    meta = dict(name=name, scheme=scheme, n=n, dt=dt, theta=theta, params=p,
                spot_max=spot_max, var_max=var_max, barrier=barrier, idx=list(map(int, O.idx)))
    return dict(meta=np.array(json.dumps(meta)), spots=spots, vars=vars_, U0=U0, V=V,
                floors=floors, price_cf=np.array(cf))


def main():
    for name in CASES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **make(name))
        print("wrote", name)


if __name__ == "__main__":
    main()
