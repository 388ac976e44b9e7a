"""One-dimensional Black-Scholes ADI engine -- oracle restatement.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates FiniteDifference/blackscholes.py:163-226 (the non-log-space
coefficients r*S and 0.5*sigma^2*S^2, the -r zero-derivative term and the
boundaries: Dirichlet 0 at S = 0, dU/dS = 1 at Smax for the first derivative,
free boundary for the second) on top of the generic operator construction and
schemes of FiniteDifference/FiniteDifferenceEngine.pyx:423-1016 (shared with
the Heston oracle).  The object also carries what the GPU engine's
``from_host_FiniteDifferenceEngine`` reads from a host engine (``grid``,
``operators``, ``simple_operators``, ``coefficients``, ``option``), so the
reference's 1-D tests (FiniteDifference/tests/test_BlackScholesOption.py:24-75)
can run against the device operators.
"""

import numpy as np

from . import banded as BO
from .engine import HestonOracle


class _Rate(object):
    def __init__(self, value):
        self.value = float(value)


class BlackScholesParams(object):
    """BlackScholesOption (blackscholes.py:19-56): spot, strike, r, variance, T."""

    def __init__(self, spot=100.0, strike=99.0, r=0.06, variance=0.04, T=1.0):
        self.spot = float(spot)
        self.strike = float(strike)
        self.r = float(r)
        self.v0 = float(variance)
        self.T = float(T)
        self.tenor = self.T
        self.interest_rate = _Rate(r)
        self.variance = _Rate(variance)
        self.barrier = None


class Grid1D(object):
    """The attributes of Grid (Grid.py:17-52) a host engine hands over."""

    def __init__(self, mesh, domain):
        self.mesh = [np.array(mesh, dtype=float)]
        self.domain = [np.array(domain, dtype=float)]
        self.shape = self.domain[-1].shape
        self.ndim = 1
        self.size = self.domain[-1].size

    def copy(self):
        return Grid1D(self.mesh[0], self.domain[-1])


class BlackScholesOracle(HestonOracle):
    """FiniteDifferenceEngineADI on one spot axis with the Black-Scholes
    coefficients; the schemes are the Heston oracle's with only axis 0."""

    def __init__(self, params, spots, **kw):
        self.option = params
        self.grid = Grid1D(spots, np.maximum(np.asarray(spots, dtype=float) - params.strike, 0))
        r, var = params.r, params.v0
        self.coefficients = {(): lambda t: -r,
                             (0,): lambda t, S: r * S,
                             (0, 0): lambda t, S: 0.5 * var * S ** 2}
        HestonOracle.__init__(self, params, spots, np.zeros(1), **kw)
        self.shape = (len(self.spots),)

    def init(self):
        """Operators are built by the constructor (FiniteDifferenceEngine.pyx
        builds them in init())."""

    def payoff(self):
        return np.maximum(self.spots - self.p.strike, 0.0)

    @property
    def idx(self):
        """blackscholes.py:228-230."""
        return int(np.argmin(np.abs(self.spots - self.p.spot)))

    def coefficient_vector(self, d, dim):
        return np.asarray(self.coefficients[d](0.0, self.spots), dtype=float)

    def boundaries(self):
        """blackscholes.py:212-216 (non-log space)."""
        return {(0,): ((0, 0.0), (1, 1.0)), (0, 0): ((0, 0.0), (None, 1.0))}

    def make_discrete_operators(self):
        """FiniteDifferenceEngine.pyx:423-548 for one axis."""
        bounds = self.boundaries()
        templates = {}
        for d in ((0,), (0, 0)):
            low, high = self._bandwidth(d, bounds[d])
            B = BO.for_vector(self.spots, derivative=len(d), force_bandwidth=(low, high),
                              axis=0)
            B.applyboundary(bounds[d])
            if bounds[d][0][0] == 0:
                B.dirichlet[0] = bounds[d][0][1]
            if bounds[d][1][0] == 0:
                B.dirichlet[1] = bounds[d][1][1]
            templates[d] = B
        self.simple_operators = templates
        self.scale_and_combine_operators()

    def scale_and_combine_operators(self):
        """FiniteDifferenceEngine.pyx:563-592 with ndim = 1: the whole -r goes
        onto the spot operator."""
        self.operators = {}
        rsplit = self.coefficients[()](0.0) / 1.0
        if self.r_split_float32:
            rsplit = float(np.float32(rsplit))
        for d in sorted(self.simple_operators):
            op = self.simple_operators[d].copy()
            op.vectorized_scale(self.coefficient_vector(d, 0))
            if 0 not in self.operators:
                self.operators[0] = op
                self.operators[0] += rsplit
            elif tuple(self.operators[0].offsets) == tuple(op.offsets):
                self.operators[0] += op
            else:
                self.operators[0] = self.operators[0] + op

    def _dims(self):
        return [0]

    def cross_term(self, V):
        return 0.0
