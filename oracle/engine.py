"""Heston ADI engine -- oracle restatement.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates FiniteDifference/FiniteDifferenceEngine.pyx (operator construction
:423-592, schemes :616-1016) with the Heston coefficient and boundary
dictionaries of FiniteDifference/heston.py:302-359, for the central scheme.
Meshes are inputs: the caller supplies the spot and variance grids.

Reference quirks and rounding-order probes are exposed as switches (both default to the behaviour
of the reference's GPU engine, which is what the product replaces):

* ``r_split_float32`` -- the CPU engine adds -r/ndim through
  ``BandedOperator.add_scalar(self, float other)`` whose Cython ``float`` is a
  C float, so the CPU engine rounds -r/2 to single precision
  (BandedOperator.pyx:713, FiniteDifferenceEngine.pyx:586).  The GPU engine
  keeps it in double (FiniteDifferenceEngineGPU.pyx:277, 693).
* ``firsts_reversed`` -- sums the explicit operators in the opposite order.
  The reference's order is Python 2 dict order (arbitrary); the difference
  between the two orders is the oracle's own rounding noise floor, which the
  parity tests use to calibrate the grid tolerance.
* ``vmax_row`` -- "reference" (default): the d2/dv2 row at v = vmax is the
  reference's free second-order backward stencil, which reaches two nodes
  back (heston.py:350-351, the -2 band the engines fold away);
  "known_derivative": the tridiagonal row of a known first derivative 0
  there, 2 (V[n-2] - V[n-1]) / h^2.  Not the reference's operator: a probe
  that shows what the reference's row does to stability on meshes whose
  first spacing is tiny (tests/test_oracle_stability.py).
* ``scale_late`` -- the explicit operators dt A and theta dt A applied as
  dt (A x + R) instead of (dt A) x + dt R: the same arithmetic associated
  differently, a rounding-order probe (tests/parity.py).
* ``implicit_residual`` -- the CPU implicit scheme keeps the boundary residual
  R in (I - dt A) x = V + dt R (FiniteDifferenceEngine.pyx:631); the GPU
  engine disables it (FiniteDifferenceEngineGPU.pyx:187-188).  Default True
  (the CPU behaviour: it enforces the Von Neumann boundary).
"""

import numpy as np

from . import banded as BO


class HestonParams(object):
    """Heston option parameters (heston.py:37-79, Option.py:17-56)."""

    def __init__(self, spot=100.0, strike=99.0, r=0.06, v0=0.04, T=1.0,
                 kappa=1.0, theta=None, sigma=0.4, rho=0.0, barrier=None):
        self.spot = float(spot)
        self.strike = float(strike)
        self.r = float(r)
        self.v0 = float(v0)
        self.T = float(T)
        self.kappa = float(kappa)
        self.theta = float(v0 if theta is None else theta)
        self.sigma = float(sigma)
        self.rho = float(rho)
        self.barrier = barrier    # None or the up-and-out level (top[1])


class HestonOracle(object):
    """FiniteDifferenceEngineADI specialised to the Heston PDE.

    ``spots`` and ``vars`` are the meshes, V has shape (len(spots), len(vars))
    with the variance axis contiguous (Grid.py:33-38).
    """

    def __init__(self, params, spots, vars, r_split_float32=False,
                 implicit_residual=True, firsts_reversed=False, tri_solver="lapack",
                 apply_order=None, vmax_row="reference", scale_late=False):
        self.p = params
        self.spots = np.array(spots, dtype=float)
        self.vars = np.array(vars, dtype=float)
        self.shape = (len(self.spots), len(self.vars))
        self.r_split_float32 = r_split_float32
        self.implicit_residual = implicit_residual
        self.firsts_reversed = firsts_reversed
        # "thomas": solve the implicit systems by per-line Thomas elimination
        # instead of LAPACK (a rounding-order probe for tests, see banded.py)
        self.tri_solver = tri_solver
        # "reversed": every operator apply sums its diagonals in reverse order
        self.apply_order = apply_order
        if vmax_row not in ("reference", "known_derivative"):
            raise ValueError("vmax_row must be 'reference' or 'known_derivative'")
        self.vmax_row = vmax_row
        self.scale_late = bool(scale_late)
        self.t = 0.0
        self.simple_operators = {}
        self.operators = {}
        self.make_discrete_operators()

    # -- initial condition / price --------------------------------------------
    def payoff(self):
        S = self.spots[:, None] + 0 * self.vars[None, :]
        return np.maximum(S - self.p.strike, 0.0)

    @property
    def idx(self):
        """heston.py:461-465 (bisect on meshes rounded to 4 decimals)."""
        from bisect import bisect_left
        ids = bisect_left(list(np.round(self.spots, 4)), round(self.p.spot, 4))
        idv = bisect_left(list(np.round(self.vars, 4)), round(self.p.v0, 4))
        return ids, idv

    # -- coefficients ------------------------------------------------------------
    def coefficient_vector(self, d, dim):
        """Coefficient of derivative ``d`` at every grid point, ordered with
        axis ``dim`` fastest (FiniteDifferenceEngine.pyx:389-418,
        heston.py:303-326)."""
        p = self.p
        S = self.spots[:, None] + np.zeros(self.shape)
        v = self.vars[None, :] + np.zeros(self.shape)
        if d == (0,):
            c = p.r * S
        elif d == (0, 0):
            c = 0.5 * v * S ** 2
        elif d == (1,):
            c = p.kappa * (p.theta - v)
            c[S == 0] = 0
        elif d == (1, 1):
            c = 0.5 * p.sigma ** 2 * v
            c[S == 0] = 0
        elif d == (0, 1):
            c = p.rho * p.sigma * S * v
        else:
            raise KeyError(d)
        if dim == 0:
            c = c.T
        return np.ascontiguousarray(c).ravel()

    def boundaries(self):
        """heston.py:336-377."""
        b = {
            (0,): ((0, 0.0), (1, 1.0)),
            (0, 0): ((0, 0.0), (None, 1.0)),
            (1,): ((None, None), (None, None)),
            (1, 1): ((1, 0.0), (None, None)),
        }
        if self.p.barrier:
            b[(0,)] = (b[(0,)][0], (0, 0.0))
            b[(0, 0)] = b[(0,)]
        if self.vmax_row == "known_derivative":
            b[(1, 1)] = (b[(1, 1)][0], (None, 0.0))
        return b

    @staticmethod
    def _bandwidth(d, bound):
        """FiniteDifferenceEngine.pyx:295-351 for the central scheme."""
        high, low = 1, -1
        order = len(d)
        if bound[0][0] is None:
            high = max(high, 1 if (order == 1 or bound[0][1] is not None) else 2)
        if bound[1][0] is None:
            low = min(low, -1 if (order == 1 or bound[1][1] is not None) else -2)
        return low, high

    # -- construction --------------------------------------------------------------
    def make_discrete_operators(self):
        """FiniteDifferenceEngine.pyx:423-548."""
        meshes = (self.spots, self.vars)
        bounds = self.boundaries()
        templates = {}
        for d in ((0,), (0, 0), (1,), (1, 1)):
            dim = d[0]
            low, high = self._bandwidth(d, bounds[d])
            B = BO.for_vector(meshes[dim], derivative=len(d),
                              force_bandwidth=(low, high), axis=dim)
            nother = self.shape[1 - dim]
            lows, highs = [], []
            for _ in range(nother):
                B.applyboundary(bounds[d])
                if bounds[d][0][0] == 0:
                    lows.append(bounds[d][0][1])
                if bounds[d][1][0] == 0:
                    highs.append(bounds[d][1][1])
            Bs = BO.flatten_blocks([B.copy() for _ in range(nother)])
            if lows:
                Bs.dirichlet[0] = tuple(lows) if len(lows) > 1 else lows[0]
            if highs:
                Bs.dirichlet[1] = tuple(highs) if len(highs) > 1 else highs[0]
            templates[d] = Bs
        templates[(0, 1)] = BO.mixed_operator(self.spots, self.vars)
        self.simple_operators = templates
        self.scale_and_combine_operators()

    def scale_and_combine_operators(self):
        """FiniteDifferenceEngine.pyx:563-592."""
        self.operators = {}
        rsplit = -self.p.r / 2.0
        if self.r_split_float32:
            rsplit = float(np.float32(rsplit))
        for d in sorted(self.simple_operators):
            op = self.simple_operators[d].copy()
            dim = op.axis
            op.vectorized_scale(self.coefficient_vector(d, dim))
            if len(set(d)) > 1:
                self.operators[d] = op
            elif dim not in self.operators:
                self.operators[dim] = op
                self.operators[dim] += rsplit
            elif tuple(self.operators[dim].offsets) == tuple(op.offsets):
                self.operators[dim] += op
            else:
                self.operators[dim] = self.operators[dim] + op

    # -- scheme building blocks ----------------------------------------------------
    def _dims(self):
        return [k for k in (0, 1) if k in self.operators]

    def _scaled(self, op, c):
        """c * op: the operator scaled (the reference), or -- with
        ``scale_late`` -- applied unscaled and its result scaled, a
        rounding-order probe: c A x + c R instead of (c A) x + c R."""
        if not self.scale_late:
            return op * c
        B = op.copy()
        B.post_scale = c
        return B

    def _firsts(self, dt):
        # The reference sums the explicit operators in Python 2 dict order
        # (FiniteDifference/FiniteDifferenceEngine.pyx:958), which is arbitrary;
        # ``firsts_reversed`` flips the order to measure that rounding noise.
        keys = list(self.operators)
        if self.firsts_reversed:
            keys = keys[::-1]
        return [self._ordered(self._scaled(self.operators[k], dt)) for k in keys]

    def _ordered(self, op):
        op.apply_order = self.apply_order
        return op

    def _les_lis(self, dt, theta, fold):
        Les = [self._ordered(self._scaled(self.operators[k], theta * dt)) for k in self._dims()]
        Lis = [(self.operators[k] * (theta * -dt)).add(1, inplace=True)
               for k in self._dims()]
        for L in Les + Lis:
            L.R = None
            L.solver = self.tri_solver
        if fold:
            for L in Les + Lis:
                if L.is_foldable():
                    L.diagonalize()
        elif self.tri_solver != "lapack":
            # the elimination probes solve tridiagonal systems: fold the
            # implicit operators (the solution is unchanged) so that the
            # variance systems are eliminated without pivoting too
            for L in Lis:
                if L.is_foldable():
                    L.diagonalize()
        return Les, Lis

    def cross_term(self, V):
        """FiniteDifferenceEngine.pyx:595-613 (numpy=False branch)."""
        if (0, 1) in self.operators:
            return self._ordered(self.operators[(0, 1)]).apply(V)
        return 0.0

    # -- schemes ---------------------------------------------------------------------
    def solve_implicit(self, n, dt, initial):
        """FiniteDifferenceEngine.pyx:616-659: cross term explicit, then
        solve the variance and the spot systems (order rolled by one)."""
        V = np.array(initial, dtype=float, copy=True)
        Lis = [(self.operators[k] * -dt).add(1, inplace=True) for k in self._dims()]
        for L in Lis:
            L.solver = self.tri_solver
            if self.tri_solver != "lapack" and L.is_foldable():
                L.diagonalize()     # eliminated without pivoting (see _les_lis)
        if not self.implicit_residual:
            for L in Lis:
                L.R = None
        Lis = Lis[1:] + Lis[:1]
        cross = self._ordered(self.operators[(0, 1)] * dt) if (0, 1) in self.operators else None
        for _ in range(int(n)):
            if cross is not None:
                V = V + cross.apply(V)
            for L in Lis:
                V = L.solve(V)
        return V

    def solve_douglas(self, n, dt, initial, theta=0.5):
        """FiniteDifferenceEngine.pyx:944-1006."""
        V = np.array(initial, dtype=float, copy=True)
        Firsts = self._firsts(dt)
        Les, Lis = self._les_lis(dt, theta, fold=True)
        for L in Firsts:
            if L.is_foldable():
                L.diagonalize()
        for _ in range(int(n)):
            Y = V.copy()
            for L in Firsts:
                Y = Y + L.apply(V)
            for Le, Li in zip(Les, Lis):
                Y = Y - Le.apply(V)
                Y = Li.solve(Y)
            V = Y
        return V

    def solve_hundsdorferverwer(self, n, dt, initial, theta=0.5):
        """FiniteDifferenceEngine.pyx:699-775."""
        V = np.array(initial, dtype=float, copy=True)
        Firsts = self._firsts(dt)
        Les, Lis = self._les_lis(dt, theta, fold=True)
        for L in Firsts:
            if L.is_foldable():
                L.diagonalize()
        for _ in range(int(n)):
            Y = V.copy()
            for L in Firsts:
                V = V + L.apply(Y)
            Z = V.copy()
            for Le, Li in zip(Les, Lis):
                Z = Z - Le.apply(Y)
                Z = Li.solve(Z)
            Y = Y - Z
            for L in Firsts:
                keep = L.R
                L.R = None
                V = V - 0.5 * L.apply(Y)
                L.R = keep
            for Le, Li in zip(Les, Lis):
                V = V - Le.apply(Z)
                V = Li.solve(V)
        return V

    def solve_craigsneyd(self, n, dt, initial, theta=0.5):
        """FiniteDifferenceEngine.pyx:842-898 (cross correction uses dt/2
        regardless of theta, :890)."""
        V = np.array(initial, dtype=float, copy=True)
        Firsts = self._firsts(dt)
        Les, Lis = self._les_lis(dt, theta, fold=False)
        for _ in range(int(n)):
            Y = V.copy()
            for L in Firsts:
                Y = Y + L.apply(V)
            Y0 = Y.copy()
            for Le, Li in zip(Les, Lis):
                Y = Y - Le.apply(V)
                Y = Li.solve(Y)
            Y = Y0 + (0.5 * dt) * self.cross_term(Y - V)
            for Le, Li in zip(Les, Lis):
                Y = Y - Le.apply(V)
                Y = Li.solve(Y)
            V = Y
        return V

    def solve_craigsneyd2(self, n, dt, initial, theta=0.5):
        """Modified Craig-Sneyd, FiniteDifferenceEngine.pyx:778-839.  The
        (1/2-theta) F(Y2-V) correction uses the Firsts with their residual
        (:829-830)."""
        V = np.array(initial, dtype=float, copy=True)
        Firsts = self._firsts(dt)
        Les, Lis = self._les_lis(dt, theta, fold=False)
        for _ in range(int(n)):
            Y = V.copy()
            for L in Firsts:
                Y = Y + L.apply(V)
            Y0 = Y.copy()
            for Le, Li in zip(Les, Lis):
                Y = Y - Le.apply(V)
                Y = Li.solve(Y)
            Y2 = Y.copy()
            Y = Y0 + theta * dt * self.cross_term(Y2 - V)
            for L in Firsts:
                Y = Y + (0.5 - theta) * L.apply(Y2 - V)
            for Le, Li in zip(Les, Lis):
                Y = Y - Le.apply(V)
                Y = Li.solve(Y)
            V = Y
        return V

    def solve_smooth(self, n, dt, initial, smoothing_steps=2, scheme=None,
                     theta=0.60):
        """FiniteDifferenceEngine.pyx:1009-1016: 2*s implicit half steps, then
        the scheme (HV by default) for the remaining n - s steps at theta=0.6."""
        if scheme is None:
            scheme = self.solve_hundsdorferverwer
        V = self.solve_implicit(smoothing_steps * 2, dt * 0.5, initial)
        return scheme(n - smoothing_steps, dt, V, theta=theta)
