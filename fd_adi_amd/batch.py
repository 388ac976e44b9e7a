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
        n = len(params)
        if not (len(spots) == len(vars) == len(dts) == n) or n == 0:
            raise ValueError("params, spots, vars and dts must have the same nonzero length")
        arr = (_lib.HestonOptionStruct * n)()
        for k, p in enumerate(params):
            a = arr[k]
            a.r, a.kappa, a.theta = float(p["r"]), float(p["kappa"]), float(p["theta"])
            a.sigma, a.rho, a.dt = float(p["sigma"]), float(p["rho"]), float(dts[k])
            a.nS, a.nv = len(spots[k]), len(vars[k])
            a.barrier = int(bool(p.get("barrier", 0)))
        self.shapes = [(len(s), len(v)) for s, v in zip(spots, vars)]
        sp = np.ascontiguousarray(np.concatenate([np.asarray(s, float) for s in spots]))
        vp = np.ascontiguousarray(np.concatenate([np.asarray(v, float) for v in vars]))
        self.stream = stream if stream is not None else _lib.default_stream()
        h = ctypes.c_void_p()
        _lib.call("fdadi_hb_create", n, ctypes.cast(arr, ctypes.c_void_p), _lib.ptr(sp),
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

    def set_shape(self, kind, threads, lines=0):
        """Launch shape of a sweep kind (0 = spot: threads per work-group and
        ``lines`` = variance columns per tile, 4 or 16; 1 = variance: lanes per
        line); 0 selects the default."""
        _lib.call("fdadi_hb_set_shape", self._h, int(kind), int(threads), int(lines))

    # launch-shape candidates per sweep kind (spot: (threads per work-group,
    # columns per tile); variance: (lanes per line, 0)); set_shape rejects
    # those too small for the grids
    SHAPES = {0: ((64, 4), (128, 4), (256, 4), (512, 4), (1024, 4), (256, 16), (512, 16)),
              1: tuple((t, 0) for t in (1, 2, 4, 8, 16, 32, 64))}

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
