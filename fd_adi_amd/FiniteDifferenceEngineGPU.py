"""ADI scheme engines on the device -- drop-in for
FiniteDifference/FiniteDifferenceEngineGPU.pyx.

Class names, constructor arguments and the scheme API are the reference's:
``solve_implicit``, ``solve_douglas``, ``solve_hundsdorferverwer``,
``solve_smooth`` (numpy in/out) and their ``*_`` device variants, plus
``solve_craigsneyd`` / ``solve_craigsneyd2`` which the reference only has on the
CPU (FiniteDifferenceEngine.pyx:778-898).

Two execution paths, both HIP:

* ``HestonFiniteDifferenceEngine`` (the Heston PDE with the reference's
  hard-coded boundaries) runs every scheme on the fused engine
  (:class:`fd_adi_amd.batch.HestonBatch`, one launch per ADI stage, all
  coefficients generated on the device).  Set ``engine.fused = False`` to run
  the operator-algebra path instead.
* Any engine whose operators were built otherwise (``from_host_...``, or
  ``fused = False``) runs the reference's operator loops on the device
  operators (BandedOperatorGPU).

Behaviour kept from the reference GPU engine: double-precision -r/ndim on the
diagonal, smoothing = 2s implicit half steps then HV at theta = 0.6, price at
the (S0, v0) node without interpolation.  Differences, all in favour of the
reference's CPU engine semantics: the implicit scheme keeps the boundary
residual R (FiniteDifference/FiniteDifferenceEngine.pyx:631 vs
FiniteDifferenceEngineGPU.pyx:187), ``solve_douglas_`` builds the operators
if they do not exist yet, and solves append the result to ``grid.domain`` so
``price`` is current.
"""

import itertools

import numpy as np

from . import _lib
from . import BandedOperatorGPU as BOG
from .Grid import Grid
from .Option import Option, BarrierOption
from .SizedArrayPtr import SizedArrayPtr
from .batch import HestonBatch
from .utils import sinh_space, exponential_space, nearest_index


def _host_scalar(x):
    if isinstance(x, SizedArrayPtr):
        return float(x.to_numpy().ravel()[0])
    return float(x)


class FiniteDifferenceEngine(object):

    def __init__(self):
        self.simple_operators = {}
        self.operators = {}
        self.t = 0
        self.n = 0
        self.barrier = 0
        self.default_scheme = 'center'
        self.default_order = 2
        self.scaling_vec = SizedArrayPtr(tag="scaling_vec")
        self.zero_derivative_coefficient = SizedArrayPtr(tag="zero_derivative_coefficient")
        self.zero_derivative_coefficient_host = None
        self.option = None
        self.grid = None
        self.gpugrid = None
        self.gpugridmesh0 = None
        self.gpugridmesh1 = None
        self.coefficients = {}
        self.fused = False

    def _fill_gpugridmesh_from_grid(self):
        i = self.grid.ndim
        if i > 0:
            self.gpugridmesh0 = SizedArrayPtr(self.grid.mesh[0])
        if i > 1:
            self.gpugridmesh1 = SizedArrayPtr(self.grid.mesh[1])

    def from_host_FiniteDifferenceEngine(self, other):
        """Move a host engine's operators and domain onto the device
        (FiniteDifferenceEngineGPU.pyx:100-128).  ``other`` must expose
        ``grid``, ``operators`` and ``simple_operators`` (host operators in the
        CPU BandedOperator layout); ``init()`` is called when present."""
        if hasattr(other, "init"):
            other.init()
        for attr in ['option', 'grid_analytical']:
            setattr(self, attr, getattr(other, attr, None))
        g = other.grid
        self.grid = g.copy() if hasattr(g, "copy") else g
        self.gpugrid = SizedArrayPtr(self.grid.domain[-1], "FDEGPU.grid")
        self._fill_gpugridmesh_from_grid()
        self.scaling_vec.alloc(self.gpugrid.size, self.scaling_vec.tag)
        self.shape = self.grid.shape
        self.ndim = self.grid.ndim
        self.coefficients = getattr(other, "coefficients", {})
        ops = dict(other.operators)
        sops = dict(getattr(other, "simple_operators", {}))
        try:
            if self.option.correlation == 0:
                ops.pop((0, 1), None)
                sops.pop((0, 1), None)
        except AttributeError:
            pass
        self.operators = {k: BOG.BandedOperator(v) for k, v in ops.items()}
        self.simple_operators = {k: BOG.BandedOperator(v) for k, v in sops.items()}
        self.fused = False
        self._prebuilt = True


class FiniteDifferenceEngineADI(FiniteDifferenceEngine):

    def __init__(self):
        FiniteDifferenceEngine.__init__(self)
        self._prebuilt = False

    def coefficient_vector(self, f, t, dim):
        """f evaluated on the grid with axis ``dim`` fastest varying
        (FiniteDifferenceEngineGPU.pyx:141-160)."""
        mesh = list(self.grid.mesh)
        m = mesh.pop(dim)
        mesh.append(m)
        args = [np.asarray(a, dtype=float) for a in
                zip(*itertools.product(*mesh))] if len(mesh) > 1 else [m]
        a = args.pop()
        args.insert(dim, a)
        ret = f(t, *args)
        if np.isscalar(ret):
            ret = np.repeat(float(ret), self.grid.size)
        return np.asarray(ret, dtype=float)

    # -- operator preparation -------------------------------------------------------
    def set_zero_derivative(self):
        """-r/ndim as a device scalar (FiniteDifferenceEngineGPU.pyx:343-361)."""
        if self.zero_derivative_coefficient_host is None:
            if isinstance(self.coefficients, dict) and () in self.coefficients:
                self.zero_derivative_coefficient_host = np.atleast_1d(
                    self.coefficients[()](self.t) / float(self.grid.ndim))
            else:
                try:
                    self.zero_derivative_coefficient_host = np.atleast_1d(
                        -self.option.interest_rate.value / self.grid.ndim)
                except AttributeError:
                    raise RuntimeError("Zero derivative coefficient has not been set.")
        self.zero_derivative_coefficient = SizedArrayPtr(
            np.asarray(self.zero_derivative_coefficient_host, dtype=float))

    def scale_and_combine_operators(self, operators=None):
        """Scale each single-derivative operator by its PDE coefficient and sum
        per dimension (FiniteDifferenceEngineGPU.pyx:193-280); on the device
        the coefficient fields come from fdadi_heston_coefficients."""
        if operators is None:
            if hasattr(self, "make_operator_templates"):
                self.make_operator_templates()
            operators = self.simple_operators
        self.operators = {}
        coeffs = self.coefficients
        on_gpu = isinstance(coeffs, list)
        for d, op in sorted(operators.items()):
            op = op.copy()
            dim = op.axis
            if not on_gpu:
                if d in coeffs:
                    op.vectorized_scale_from_host(self.coefficient_vector(coeffs[d], self.t, dim))
            elif d in coeffs:
                which = {(0,): 0, (0, 0): 1, (1,): 2, (1, 1): 3, (0, 1): 4}[d]
                o = self.option
                p = np.array([o.interest_rate.value, o.variance.reversion, o.variance.mean,
                              o.variance.volatility, o.correlation], dtype=float)
                _lib.call("fdadi_heston_coefficients", which, self.gpugridmesh0.p,
                          self.gpugridmesh0.size, self.gpugridmesh1.p, self.gpugridmesh1.size,
                          _lib.ptr(p), self.scaling_vec.p, _lib.default_stream())
                op.vectorized_scale(self.scaling_vec)
            if len(set(d)) > 1:
                self.operators[d] = op
            elif dim not in self.operators:
                self.operators[dim] = op
                if () in coeffs:
                    if not on_gpu:
                        self.operators[dim] += coeffs[()](self.t) / float(self.grid.ndim)
                    elif self.zero_derivative_coefficient.size == 0:
                        raise AssertionError("Zero derviative has not been set.")
                    else:
                        self.operators[dim].add_scalar(self.zero_derivative_coefficient, self.n)
            else:
                self.operators[dim] += op
        _lib.sync()
        return self.scaling_vec

    def preprocess_operators(self, dt, theta):
        """Firsts = dt*A, Les = theta*dt*A, Lis = I - theta*dt*A (residual
        off for Les/Lis) (FiniteDifferenceEngineGPU.pyx:163-190)."""
        if self.zero_derivative_coefficient.size == 0:
            self.set_zero_derivative()
        if not self._prebuilt:
            self.scale_and_combine_operators()
        dt = dt if isinstance(dt, SizedArrayPtr) else SizedArrayPtr(np.atleast_1d(float(dt)))
        theta = (theta if isinstance(theta, SizedArrayPtr)
                 else SizedArrayPtr(np.atleast_1d(float(theta))))
        withdt = {k: (o * dt) for k, o in self.operators.items()}
        Firsts = list(withdt.values())
        dims = sorted(k for k in withdt if not isinstance(k, tuple))
        Les = [withdt[k] * theta for k in dims]
        neg = -theta
        Lis = [(withdt[k] * neg).add(1, inplace=True) for k in dims]
        for L in itertools.chain(Les, Lis):
            L.enable_residual(False)
        return Firsts, Les, Lis

    # -- numpy wrappers -----------------------------------------------------------------
    def _run(self, fn, n, dt, initial, **kw):
        n = int(n)
        V = SizedArrayPtr(initial) if initial is not None else self.gpugrid
        fn(n, dt, V, **kw)
        ret = V.to_numpy()
        if self.grid is not None:
            self.grid.domain.append(ret.copy())
        return ret

    def solve_implicit(self, n, dt, initial=None, callback=None, numpy=False):
        if callback or numpy:
            raise NotImplementedError("Callbacks and Numpy not available for GPU solver.")
        return self._run(self.solve_implicit_, n, dt, initial)

    def solve_douglas(self, n, dt, initial=None, theta=0.5, callback=None, numpy=False):
        if callback or numpy:
            raise NotImplementedError("Callbacks and Numpy not available for GPU solver.")
        return self._run(self.solve_douglas_, n, dt, initial, theta=theta)

    def solve_hundsdorferverwer(self, n, dt, initial=None, theta=0.5, callback=None,
                                numpy=False):
        if callback or numpy:
            raise NotImplementedError("Callbacks and Numpy not available for GPU solver.")
        return self._run(self.solve_hundsdorferverwer_, n, dt, initial, theta=theta)

    def solve_craigsneyd(self, n, dt, initial=None, theta=0.5, callback=None, numpy=False):
        if callback or numpy:
            raise NotImplementedError("Callbacks and Numpy not available for GPU solver.")
        return self._run(self.solve_craigsneyd_, n, dt, initial, theta=theta)

    def solve_craigsneyd2(self, n, dt, initial=None, theta=0.5, callback=None, numpy=False):
        if callback or numpy:
            raise NotImplementedError("Callbacks and Numpy not available for GPU solver.")
        return self._run(self.solve_craigsneyd2_, n, dt, initial, theta=theta)

    def solve_smooth(self, n, dt, initial=None, callback=None, smoothing_steps=2,
                     scheme=None):
        if scheme:
            raise NotImplementedError("Changing smoothing schemes not supported on GPU.")
        return self._run(self.solve_smooth_, n, dt, initial, smoothing_steps=smoothing_steps)

    # -- device schemes (operator-algebra path) -----------------------------------------
    def _fused_ok(self):
        return False

    def solve_implicit_(self, n, dt, V, callback=None, numpy=False):
        """Fully implicit: cross term explicit, then variance and spot solves.
        ``dt`` may be a device scalar (SizedArrayPtr), as in the reference
        (FiniteDifferenceEngineGPU.pyx:379)."""
        if callback or numpy:
            raise NotImplementedError("Callbacks and Numpy not available for GPU solver.")
        if self._fused_ok():
            return self._fused(V, "implicit", n, dt, 1.0)
        f, l, Lis = self.preprocess_operators(dt, 1.0)
        del f, l
        for L in Lis:
            L.enable_residual(True)  # (I - dt A) x = V + dt R, as the CPU engine
        Lis = Lis[1:] + Lis[:1]
        cross = None
        if (0, 1) in self.operators:
            dtp = dt if isinstance(dt, SizedArrayPtr) else SizedArrayPtr(np.atleast_1d(float(dt)))
            cross = self.operators[(0, 1)] * dtp
        for o in Lis:
            if o.is_foldable():
                o.diagonalize()
        U = SizedArrayPtr().alloc(V.size)
        for _ in range(int(n)):
            if cross is not None:
                U.copy_from(V)
                cross.apply_(U, overwrite=True)
                V.pluseq(U)
            for L in Lis:
                L.solve_(V, overwrite=True)
        _lib.sync()

    def solve_douglas_(self, n, dt, V, theta=0.5):
        if self._fused_ok():
            return self._fused(V, "douglas", n, _host_scalar(dt), theta)
        if not self.operators:
            self.preprocess_operators(1.0, 1.0)  # builds self.operators
        dt = _host_scalar(dt)
        ops = self.operators
        dims = sorted(k for k in ops if not isinstance(k, tuple))
        Firsts = [o.mul_scalar_from_host(dt) for o in ops.values()]
        Les = [ops[k].mul_scalar_from_host(theta * dt) for k in dims]
        Lis = [ops[k].mul_scalar_from_host(theta * -dt).add(1, inplace=True) for k in dims]
        for L in itertools.chain(Les, Lis):
            L.enable_residual(False)
        for L in itertools.chain(Les, Lis, Firsts):
            if L.is_foldable():
                L.diagonalize()
        Y = V.copy(True)
        X = SizedArrayPtr().alloc(V.size)
        for _ in range(int(n)):
            for L in Firsts:
                X.copy_from(Y)
                L.apply_(X, overwrite=True)
                V.pluseq(X)
            for Le, Li in zip(Les, Lis):
                X.copy_from(Y)
                Le.apply_(X, overwrite=True)
                V.minuseq(X)
                Li.solve_(V, overwrite=True)
            Y.copy_from(V)
        _lib.sync()

    def solve_hundsdorferverwer_(self, n, dt, V, theta=0.5):
        """``dt`` and ``theta`` may be device scalars (SizedArrayPtr), as in
        the reference (FiniteDifferenceEngineGPU.pyx:502)."""
        if self._fused_ok():
            return self._fused(V, "hv", n, dt, theta)
        Firsts, Les, Lis = self.preprocess_operators(dt, theta)
        for L in itertools.chain(Les, Lis, Firsts):
            if L.is_foldable():
                L.diagonalize()
        X = SizedArrayPtr().alloc(V.size)
        Y = SizedArrayPtr().alloc(V.size)
        Z = SizedArrayPtr().alloc(V.size)
        for _ in range(int(n)):
            Y.copy_from(V)
            for L in Firsts:
                X.copy_from(Y)
                L.apply_(X, overwrite=True)
                V.pluseq(X)
            Z.copy_from(V)
            for Le, Li in zip(Les, Lis):
                X.copy_from(Y)
                Le.apply_(X, overwrite=True)
                Z.minuseq(X)
                Li.solve_(Z, overwrite=True)
            Y.minuseq_over2(Z)
            for L in Firsts:
                L.enable_residual(False)
                X.copy_from(Y)
                L.apply_(X, overwrite=True)
                V.minuseq(X)
                L.enable_residual(True)
            for Le, Li in zip(Les, Lis):
                X.copy_from(Z)
                Le.apply_(X, overwrite=True)
                V.minuseq(X)
                Li.solve_(V, overwrite=True)
        _lib.sync()

    def _cs_common(self, n, dt, V, theta, modified):
        Firsts, Les, Lis = self.preprocess_operators(dt, theta)
        dt_h, th = _host_scalar(dt), _host_scalar(theta)
        cross = self.operators.get((0, 1))
        X = SizedArrayPtr().alloc(V.size)
        Y = SizedArrayPtr().alloc(V.size)
        Y0 = SizedArrayPtr().alloc(V.size)
        D = SizedArrayPtr().alloc(V.size)
        for _ in range(int(n)):
            Y.copy_from(V)
            for L in Firsts:
                X.copy_from(V)
                L.apply_(X, overwrite=True)
                Y.pluseq(X)
            Y0.copy_from(Y)
            for Le, Li in zip(Les, Lis):
                X.copy_from(V)
                Le.apply_(X, overwrite=True)
                Y.minuseq(X)
                Li.solve_(Y, overwrite=True)
            D.copy_from(Y)
            D.minuseq(V)              # D = Y2 - V
            Y.copy_from(Y0)
            if cross is not None:
                X.copy_from(D)
                cross.apply_(X, overwrite=True)
                X.timeseq_scalar_from_host(th * dt_h if modified else 0.5 * dt_h)
                Y.pluseq(X)
            if modified:
                for L in Firsts:      # (1/2 - theta) F(Y2 - V), residual included
                    X.copy_from(D)
                    L.apply_(X, overwrite=True)
                    X.timeseq_scalar_from_host(0.5 - th)
                    Y.pluseq(X)
            for Le, Li in zip(Les, Lis):
                X.copy_from(V)
                Le.apply_(X, overwrite=True)
                Y.minuseq(X)
                Li.solve_(Y, overwrite=True)
            V.copy_from(Y)
        _lib.sync()

    def solve_craigsneyd_(self, n, dt, V, theta=0.5):
        """Craig-Sneyd (FiniteDifference/FiniteDifferenceEngine.pyx:842-898):
        the cross correction uses dt/2 whatever theta is."""
        if self._fused_ok():
            return self._fused(V, "cs", n, _host_scalar(dt), _host_scalar(theta))
        return self._cs_common(n, dt, V, theta, modified=False)

    def solve_craigsneyd2_(self, n, dt, V, theta=0.5):
        """Modified Craig-Sneyd (FiniteDifference/FiniteDifferenceEngine.pyx:778-839)."""
        if self._fused_ok():
            return self._fused(V, "mcs", n, _host_scalar(dt), _host_scalar(theta))
        return self._cs_common(n, dt, V, theta, modified=True)

    def solve_smooth_(self, n, dt, V, smoothing_steps=2):
        """2*s implicit half steps, then HV at theta = 0.6 to completion
        (FiniteDifferenceEngineGPU.pyx:582-593); a device ``dt`` is halved on
        the device, as the reference does."""
        if isinstance(dt, SizedArrayPtr):
            half = SizedArrayPtr().alloc(1)
            half.copy_from(dt)
            half.timeseq_scalar_from_host(0.5)
        else:
            half = 0.5 * float(dt)
        self.solve_implicit_(smoothing_steps * 2, half, V)
        self.solve_hundsdorferverwer_(int(n) - smoothing_steps, dt, V, 0.6)

    def _dummy(self):
        """One Douglas step from arange(N) (FiniteDifferenceEngineGPU.pyx:283-340)."""
        n, dt, theta = 1, 0.01, 0.5
        initial = np.arange(self.shape[0] * self.shape[1], dtype=float).reshape(self.shape)
        ops = self.operators
        dims = sorted(k for k in ops if not isinstance(k, tuple))
        Firsts = [o.mul_scalar_from_host(dt) for d, o in ops.items()]
        Les = [ops[k].mul_scalar_from_host(theta * dt) for k in dims]
        Lis = [ops[k].mul_scalar_from_host(-dt * theta).add(1, inplace=True) for k in dims]
        for L in itertools.chain(Les, Lis):
            L.enable_residual(False)
        V = SizedArrayPtr(initial)
        Orig = V.copy(True)
        Y = V.copy(True)
        X = V.copy(True)
        for L in itertools.chain(Les, Lis, Firsts):
            if L.is_foldable():
                L.diagonalize()
        for _ in range(n):
            Y.copy_from(V)
            for L in Firsts:
                X.copy_from(Y)
                L.apply_(X, overwrite=True)
                V.pluseq(X)
            for Le, Li in zip(Les, Lis):
                X.copy_from(Y)
                Le.apply_(X, overwrite=True)
                V.minuseq(X)
                Li.solve_(V, overwrite=True)
        return Firsts, Les, Lis, Orig.to_numpy(), Y.to_numpy(), V.to_numpy()


class HestonFiniteDifferenceEngine(FiniteDifferenceEngineADI):
    """Heston engine with the reference's hard-coded coefficients and
    boundaries (FiniteDifferenceEngineGPU.pyx:596-738)."""

    def __init__(self, option, grid=None, spot_max=1500.0, spot_min=0.0, spots=None,
                 vars=None, var_max=10.0, nspots=100, nvols=100, spotdensity=7.0,
                 varexp=4.0, force_exact=True, flip_idx_var=False, flip_idx_spot=False,
                 schemes=None, coefficients=None, boundaries=None, cache=True,
                 verbose=True, force_bandwidth=None):
        FiniteDifferenceEngineADI.__init__(self)
        if schemes is not None or flip_idx_var or flip_idx_spot:
            raise NotImplementedError("Only central differencing supported on GPU")
        self.coefficients = [(0,), (0, 0), (1,), (1, 1), (0, 1), ()]
        self.cache = cache
        if not isinstance(option, Option) and not hasattr(option, "variance"):
            raise AssertionError("option must be an Option")
        self.option = option
        if isinstance(option, BarrierOption) or getattr(option, "top", None) is not None \
                or getattr(option, "bottom", None) is not None:
            top = getattr(option, "top", None)
            if top:
                if top[0]:
                    raise NotImplementedError("Knockin barriers are not supported.")
                spot_max = top[1]
                if grid:
                    assert np.allclose(spot_max, max(grid.mesh[0]))
            if getattr(option, "bottom", None):
                raise NotImplementedError("Bottom barriers are not supported for GPU creation.")
        if grid:
            # The grid's own mesh arrays: the S0 snap below moves grid.mesh[0]
            # too (FiniteDifferenceEngineGPU.pyx:659-661).
            self.spots = grid.mesh[0]
            self.vars = grid.mesh[1]
        else:
            if vars is None:
                vars = exponential_space(0.00, option.variance.value, var_max, varexp, nvols,
                                         force_exact=force_exact)
            self.vars = np.asarray(vars, dtype=float)
            if spots is None:
                top = getattr(option, "top", None)
                if top and not top[0]:
                    spots = exponential_space(0.00, option.strike, spot_max, 1.0 / 3, nspots,
                                              force_exact=force_exact)
                    self.barrier = top[1]
                else:
                    spots = sinh_space(option.strike - spot_min, spot_max - spot_min,
                                       spotdensity, nspots, force_exact=force_exact) + spot_min
            K = option.strike
            # Payoff on the mesh as generated, before S0 is snapped onto it
            # (FiniteDifferenceEngineGPU.pyx:682).
            grid = Grid([np.asarray(spots, dtype=float), self.vars],
                        initializer=lambda *x: np.maximum(x[0] - K, 0))
            self.spots = grid.mesh[0]
            self.vars = grid.mesh[1]
        # S0 onto its nearest node, in place: self.spots is grid.mesh[0], so the
        # grid's mesh moves with it while the payoff keeps the unsnapped node
        # (FiniteDifferenceEngineGPU.pyx:685-686).
        self.spots[np.argmin(np.abs(self.spots - option.spot))] = option.spot
        if getattr(option, "top", None) and not option.top[0]:
            self.barrier = option.top[1]
        self.grid = grid
        self.shape = grid.shape
        self.ndim = grid.ndim
        self.gpugrid = SizedArrayPtr(self.grid.domain[-1], "FDEGPU.grid")
        self._fill_gpugridmesh_from_grid()
        self.scaling_vec.alloc(self.gpugrid.size, self.scaling_vec.tag)
        self.zero_derivative_coefficient_host = np.atleast_1d(
            -self.option.interest_rate.value / self.grid.ndim)
        self.fused = True
        self._batch = None

    def make_operator_templates(self):
        """Single-derivative Heston operators (FiniteDifferenceEngineGPU.pyx:697-724)."""
        m0 = self.grid.mesh[0]
        m1 = self.grid.mesh[1]
        self.zero_derivative_coefficient_host = np.atleast_1d(
            -self.option.interest_rate.value / self.grid.ndim)
        self.simple_operators[(0,)] = BOG.for_vector(m0, m1.size, 1, 0, self.barrier)
        self.simple_operators[(0,)].has_low_dirichlet = True
        if self.barrier:
            self.simple_operators[(0,)].has_high_dirichlet = True
        self.simple_operators[(0, 0)] = BOG.for_vector(m0, m1.size, 2, 0, self.barrier)
        self.simple_operators[(0, 0)].has_low_dirichlet = True
        if self.barrier:
            self.simple_operators[(0, 0)].has_high_dirichlet = True
        self.simple_operators[(1,)] = BOG.for_vector(m1, m0.size, 1, 1, self.barrier)
        self.simple_operators[(1, 1)] = BOG.for_vector(m1, m0.size, 2, 1, self.barrier)
        if getattr(self.option, "correlation", 0) != 0:
            self.simple_operators[(0, 1)] = BOG.mixed_for_vector(m0, m1)

    # -- fused path ---------------------------------------------------------------------
    def _fused_ok(self):
        return bool(self.fused)

    def batch(self):
        """The one-option fused engine for this grid (unit base step; each
        solve scales it to its dt exactly)."""
        if self._batch is None:
            o = self.option
            params = [dict(r=o.interest_rate.value, kappa=o.variance.reversion,
                           theta=o.variance.mean, sigma=o.variance.volatility,
                           rho=getattr(o, "correlation", 0.0), barrier=bool(self.barrier))]
            self._batch = HestonBatch(params, [self.spots], [self.vars], [1.0])
        return self._batch

    def _fused(self, V, scheme, n, dt, theta):
        """Run on the one-option fused engine (unit base step, so dt is the
        multiplier); with a device dt or theta the scheme is configured from
        device memory and nothing is read back."""
        hb = self.batch()
        view = hb.state_view(0)
        view.copy_from(V)
        if isinstance(dt, SizedArrayPtr) or isinstance(theta, SizedArrayPtr):
            dev = [x if isinstance(x, SizedArrayPtr) else SizedArrayPtr(np.atleast_1d(float(x)))
                   for x in (theta, dt)]
            hb.configure_dev(scheme, dev[0], dev[1])
        else:
            hb.configure(scheme, float(theta), float(dt))
        hb.step(int(n))
        V.copy_from(view)
        _lib.sync()

    # -- price ------------------------------------------------------------------------
    @property
    def idx(self):
        """Indices of (S0, v0) on the meshes (FiniteDifferenceEngineGPU.pyx:727-734)."""
        return (nearest_index(self.spots, self.option.spot),
                nearest_index(self.vars, self.option.variance.value))

    @property
    def price(self):
        return self.grid.domain[-1][self.idx]
