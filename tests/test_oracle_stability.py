"""Why the 2048 x 1024 headline solve diverges (CPU only, the oracle).

The power-law variance mesh of that configuration (1024 points,
exponential_space(0, 0.04, 0.2, 4, 1024): first spacing 1.8e-13) together
with the reference's own v = vmax row of d2/dv2 (the free second-order
backward stencil, heston.py:350-351) and the cross term (rho != 0) makes the
Hundsdorfer-Verwer scheme unstable: the oracle's max|V| leaves the payoff's
range after ~150 steps and passes 1e10 by step 300 at 512 x 1024.  Removing
either ingredient -- rho = 0, or the tridiagonal known-derivative row at
vmax instead of the reference's -- keeps the same solve at the payoff's
scale.  So the divergence is the reference algorithm's, not the GPU
engine's (tests/test_gpu_batch.py::test_reference_instability_reproduced
shows the GPU engine reproducing the oracle on it).
"""

import numpy as np

from fd_adi_amd.utils import sinh_space, exponential_space
from tests.oracle_pool import solve_many

NS, NV, STEPS, DT = 512, 1024, 300, 1 / 2000.0


def test_vmax_row_and_cross_term_make_hv_diverge():
    spots = sinh_space(100.0, 500.0, 7.0, NS)
    spots[np.argmin(np.abs(spots - 100.0))] = 100.0
    vars_ = exponential_space(0.0, 0.04, 0.2, 4.0, NV)
    assert vars_[1] - vars_[0] < 1e-12
    pf = dict(spot=100, strike=100, r=0.06, v0=0.04, T=1.0, kappa=1.0, theta=0.04, sigma=0.2)
    jobs = [(dict(pf, rho=-0.7), spots, vars_, {}, "solve_hundsdorferverwer",
             (STEPS, DT, None, 0.5)),
            (dict(pf, rho=0.0), spots, vars_, {}, "solve_hundsdorferverwer",
             (STEPS, DT, None, 0.5)),
            (dict(pf, rho=-0.7), spots, vars_, dict(vmax_row="known_derivative"),
             "solve_hundsdorferverwer", (STEPS, DT, None, 0.5))]
    ref, no_cross, tri_row = (float(np.max(np.abs(V))) for V in solve_many(jobs))
    print("max|V| after %d HV steps: reference %.3e, rho = 0 %.3e, known-derivative "
          "vmax row %.3e" % (STEPS, ref, no_cross, tri_row))
    payoff_max = spots[-1] - 100.0
    assert ref > 1e10
    assert no_cross < 1.01 * payoff_max and tri_row < 1.01 * payoff_max
