"""Batched Heston ADI on the device (the fused engine behind every scheme).

``HestonBatch`` holds many options with uneven grids in one device arena and
advances them together: one kernel launch per ADI stage covers the whole
batch (fdadi_hb_* in include/fdadi.h).  A single-option engine is a batch of
one.  Per-option semantics are those of the reference schemes
(FiniteDifference/FiniteDifferenceEngine.pyx:616-1016).
"""

import ctypes

import numpy as np

from . import _lib
from .SizedArrayPtr import SizedArrayPtr

# fdadi_heston_option (include/fdadi.h) as a NumPy record
_OPTION_DTYPE = np.dtype([("r", "f8"), ("kappa", "f8"), ("theta", "f8"), ("sigma", "f8"),
                          ("rho", "f8"), ("dt", "f8"), ("nS", "i8"), ("nv", "i8"),
                          ("barrier", "i4"), ("pad", "i4")])
assert _OPTION_DTYPE.itemsize == ctypes.sizeof(_lib.HestonOptionStruct)

SCHEMES = {
    "implicit": _lib.SCHEME_IMPLICIT, "i": _lib.SCHEME_IMPLICIT,
    "douglas": _lib.SCHEME_DOUGLAS, "d": _lib.SCHEME_DOUGLAS,
    "craigsneyd": _lib.SCHEME_CS, "cs": _lib.SCHEME_CS,
    "craigsneyd2": _lib.SCHEME_MCS, "mcs": _lib.SCHEME_MCS,
    "hundsdorferverwer": _lib.SCHEME_HV, "hv": _lib.SCHEME_HV,
}


class HestonBatch(object):
    """Options k = 0..n-1 with parameters ``params[k]`` (dict with r, kappa,
    theta, sigma, rho; optional barrier), meshes ``spots[k]``/``vars[k]`` and
    step ``dts[k]``."""

    def __init__(self, params, spots, vars, dts, stream=None):
        arr, sp, vp = self._pack(params, spots, vars, dts)
        n = len(arr)
        self.shapes = [(int(a), int(b)) for a, b in zip(arr["nS"], arr["nv"])]
        self.stream = stream if stream is not None else _lib.default_stream()
        h = ctypes.c_void_p()
        _lib.call("fdadi_hb_create", n, arr.ctypes.data_as(ctypes.c_void_p), _lib.ptr(sp),
                  _lib.ptr(vp), self.stream, ctypes.byref(h))
        self._h = h
        self.n = n
        off = np.zeros(n, dtype=np.int64)
        _lib.call("fdadi_hb_offsets", h, off.ctypes.data_as(ctypes.c_void_p))
        self.offsets = off
        p, tot = ctypes.c_void_p(), _lib.c_i64()
        _lib.call("fdadi_hb_state", h, ctypes.byref(p), ctypes.byref(tot))
        self._state = p.value
        self.total = tot.value
        self.scheme = None

    @staticmethod
    def _pack(params, spots, vars, dts):
        """fdadi_heston_option records and the packed meshes.  ``params`` is a
        list of dicts or one dict of per-option arrays; ``spots``/``vars`` are
        lists of meshes or 2-D arrays with one mesh per row."""
        n = len(dts)
        if isinstance(params, dict):
            cols = {k: np.broadcast_to(np.asarray(v, float), (n,)) for k, v in params.items()}
        else:
            if len(params) != n:
                raise ValueError("params and dts must have the same length")
            cols = {k: np.array([float(p.get(k, 0.0)) for p in params])
                    for k in ("r", "kappa", "theta", "sigma", "rho", "barrier")}
        if not (len(spots) == len(vars) == n) or n == 0:
            raise ValueError("params, spots, vars and dts must have the same nonzero length")
        arr = np.zeros(n, dtype=_OPTION_DTYPE)
        for k in ("r", "kappa", "theta", "sigma", "rho"):
            arr[k] = cols[k]
        arr["dt"] = np.asarray(dts, float)
        arr["barrier"] = (np.asarray(cols.get("barrier", np.zeros(n))) != 0).astype(np.int32)
        if isinstance(spots, np.ndarray) and spots.ndim == 2:
            arr["nS"], sp = spots.shape[1], np.ascontiguousarray(spots, dtype=float).ravel()
        else:
            arr["nS"] = [len(s) for s in spots]
            sp = np.ascontiguousarray(np.concatenate([np.asarray(s, float) for s in spots]))
        if isinstance(vars, np.ndarray) and vars.ndim == 2:
            arr["nv"], vp = vars.shape[1], np.ascontiguousarray(vars, dtype=float).ravel()
        else:
            arr["nv"] = [len(v) for v in vars]
            vp = np.ascontiguousarray(np.concatenate([np.asarray(v, float) for v in vars]))
        return arr, sp, vp

    def update(self, params, spots, vars, dts):
        """New parameters, meshes and steps for the same options (same count
        and grid shapes); a configured scheme and a captured graph are kept."""
        arr, sp, vp = self._pack(params, spots, vars, dts)
        _lib.call("fdadi_hb_update", self._h, len(arr), arr.ctypes.data_as(ctypes.c_void_p),
                  _lib.ptr(sp), _lib.ptr(vp), self.stream)

    def call_payoff(self, strikes):
        """Initial state max(S - K_k, 0) on every option's grid (``strikes``:
        a device array of n strikes)."""
        _lib.call("fdadi_hb_call_payoff", self._h, strikes.p, self.stream)

    def read_bicubic(self, S0, v0, out):
        """Every option's state read at (S0[k], v0[k]) by bicubic Lagrange
        interpolation into ``out`` (device arrays of n values)."""
        _lib.call("fdadi_hb_read_bicubic", self._h, S0.p, v0.p, out.p, self.stream)

    def __del__(self):
        try:
            if self._h:
                _lib.lib.fdadi_hb_destroy(self._h)
        except Exception:
            pass
        self._h = None

    # -- configuration ---------------------------------------------------------------
    def configure(self, scheme, theta=0.5, dt_scale=1.0):
        code = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        _lib.call("fdadi_hb_configure", self._h, code, float(theta), float(dt_scale),
                  self.stream)
        self.scheme = code
        return self

    def configure_dev(self, scheme, theta, dt_scale):
        """configure() with theta and the dt multiplier in device memory
        (SizedArrayPtr of one value each); nothing is read back."""
        code = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        _lib.call("fdadi_hb_configure_dev", self._h, code, theta.p, dt_scale.p, self.stream)
        self.scheme = code
        return self

    def set_shape(self, kind, threads, lines=0):
        """Launch shape of a sweep kind (0 = spot: threads per work-group and
        ``lines`` = variance columns per tile, 4, 8, 16 or 32; 1 = variance: lanes
        per line); 0 selects the default, -1 a shape per option from its line
        length (options of one shape share a launch)."""
        _lib.call("fdadi_hb_set_shape", self._h, int(kind), int(threads), int(lines))

    # launch-shape candidates per sweep kind (spot: (threads per work-group,
    # columns per tile); variance: (lanes per line, 0)); set_shape rejects
    # those too small for the grids
    SHAPES = {0: ((64, 4), (128, 4), (256, 4), (512, 4), (256, 8), (512, 8), (256, 16),
                  (512, 16), (512, 32), (-1, 4), (-1, 8), (-1, 16)),
              1: tuple((t, 0) for t in (1, 2, 4, 8, 16, 32, 64, -1))}

    def autotune(self, steps=2):
        """Pick the fastest launch shape of each sweep kind for the configured
        scheme: every valid candidate is timed over ``steps`` steps with HIP
        events (the stage slots of that kind summed); the state is saved and
        restored.  Returns {kind: (threads, lines)} as set."""
        if self.scheme is None:
            raise RuntimeError("autotune: configure the engine first")
        # stage slots of each kind in a step (fdadi_hb_step's launch order)
        order = {0: (1, 0), 1: (0, 1)}.get(self.scheme, (0, 1, 0, 1))
        saved = self.state_view().copy(True)
        best = {}
        for kind in (0, 1):
            timings = []
            for c in self.SHAPES[kind]:
                try:
                    self.set_shape(kind, *c)
                except (ValueError, NotImplementedError):
                    continue
                self.step_timed(1)
                st = self.step_timed(steps)
                timings.append((sum(t for t, k in zip(st, order) if k == kind), c))
            best[kind] = min(timings)[1]
            self.set_shape(kind, *best[kind])
        self.state_view().copy_from(saved)
        self.sync()
        return best

    def get_shape(self, kind):
        t, l = ctypes.c_int(), ctypes.c_int()
        _lib.call("fdadi_hb_get_shape", self._h, int(kind), ctypes.byref(t), ctypes.byref(l))
        return t.value, l.value

    def launches_per_step(self):
        n = ctypes.c_int()
        _lib.call("fdadi_hb_launches_per_step", self._h, ctypes.byref(n))
        return n.value

    # -- state ---------------------------------------------------------------------------
    def state_view(self, k=None):
        """SizedArrayPtr view of option k's state (or of the whole packed state)."""
        if k is None:
            return SizedArrayPtr.from_device(self._state, (self.total,), owner_ref=self)
        return SizedArrayPtr.from_device(self._state + 8 * int(self.offsets[k]),
                                         self.shapes[k], owner_ref=self)

    def set_state(self, states):
        """Upload per-option initial states (list of (nS, nv) arrays) or a packed vector."""
        if isinstance(states, np.ndarray) and states.ndim == 1 and states.size == self.total:
            packed = np.ascontiguousarray(states, dtype=float)
        else:
            packed = np.ascontiguousarray(np.concatenate(
                [np.asarray(s, float).reshape(-1) for s in states]))
            if packed.size != self.total:
                raise ValueError("state size %d != %d" % (packed.size, self.total))
        _lib.call("fdadi_array_h2d", ctypes.c_void_p(self._state), _lib.ptr(packed),
                  self.total, self.stream)

    def get_packed(self):
        out = np.empty(self.total)
        _lib.call("fdadi_array_d2h", _lib.ptr(out), ctypes.c_void_p(self._state), self.total,
                  self.stream)
        return out

    def get_state(self):
        packed = self.get_packed()
        return [packed[o:o + a * b].reshape(a, b)
                for o, (a, b) in zip(self.offsets, self.shapes)]

    # -- stepping ------------------------------------------------------------------------
    def step(self, nsteps):
        if self.scheme is None:
            raise RuntimeError("HestonBatch.step: call configure() first")
        _lib.call("fdadi_hb_step", self._h, int(nsteps), self.stream)

    def step_timed(self, nsteps):
        """Step with HIP events around every stage; returns the summed device
        milliseconds per stage slot (launch order within a step)."""
        out = np.zeros(8)
        n = ctypes.c_int32()
        _lib.call("fdadi_hb_step_timed", self._h, int(nsteps), self.stream, _lib.ptr(out), 8,
                  ctypes.byref(n))
        return out[:n.value].copy()

    def capture(self, nsteps):
        _lib.call("fdadi_hb_capture", self._h, int(nsteps), self.stream)

    def replay(self):
        _lib.call("fdadi_hb_replay", self._h, self.stream)

    def capture_solve(self, scheme, n, theta=0.5, smoothing_steps=0):
        """Capture ``solve(scheme, n, theta, smoothing_steps)`` as one graph
        (``replay`` runs it from the current state): with ``smoothing_steps``
        = s, 2s implicit half steps, then n - s steps of ``scheme``.  Leaves
        the engine configured for ``scheme``."""
        n, s = int(n), int(smoothing_steps)
        if n - s <= 0:
            raise ValueError("capture_solve: n must exceed the smoothing steps")
        code = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        _lib.call("fdadi_hb_capture_solve", self._h, code, float(theta), s, n - s, self.stream)
        self.scheme = code

    def sync(self):
        _lib.call("fdadi_stream_sync", self.stream)

    def solve(self, scheme, n, theta=0.5, smoothing_steps=0, smooth_theta=None):
        """Run ``n`` steps of ``scheme``; with ``smoothing_steps`` = s, first
        2s implicit half steps then n - s steps of the scheme
        (FiniteDifference/FiniteDifferenceEngine.pyx:1009-1016)."""
        n = int(n)
        if smoothing_steps:
            self.configure("implicit", 1.0, 0.5)
            self.step(2 * smoothing_steps)
            n -= smoothing_steps
            if smooth_theta is not None:
                theta = smooth_theta
        self.configure(scheme, theta, 1.0)
        self.step(n)
