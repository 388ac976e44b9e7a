"""GPU schemes against the committed golden fixtures (tests/golden/*.npz).

Each fixture carries its inputs (parameters, meshes, payoff, steps) and the
oracle's grid; the GPU engine is built on the stored meshes and must meet the
parity bar of tests/parity.py against it (relative price error at (S0, v0)
<= 1e-12; grid within max(1e-10, 4 x the stored rounding-order floor)), on
both the fused path and the operator-algebra path.
"""

import pytest

from tests.golden_cases import names, load, solve
from tests.parity import check

pytestmark = pytest.mark.gpu


def engine(c):
    from fd_adi_amd.FiniteDifferenceEngineGPU import HestonFiniteDifferenceEngine
    from fd_adi_amd.Option import HestonOption, HestonBarrierOption
    p = c.p
    kw = dict(spot=p["spot"], strike=p["strike"], interest_rate=p["r"], variance=p["v0"],
              tenor=p["T"], mean_reversion=p["kappa"], mean_variance=p["theta"],
              vol_of_variance=p["sigma"], correlation=p["rho"])
    opt = HestonBarrierOption(top=(False, c.barrier), **kw) if c.barrier else HestonOption(**kw)
    return HestonFiniteDifferenceEngine(opt, spots=c.spots, vars=c.vars, verbose=False)


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "operators"])
@pytest.mark.parametrize("name", names())
def test_gpu_matches_golden(gpu, name, fused):
    c = load(name)
    G = engine(c)
    assert tuple(G.idx) == c.idx
    G.fused = fused
    V = solve(G, c, c.U0.copy())
    print(name, "fused" if fused else "operators", end=": ")
    check(V, c.V, c.floor, c.idx)
