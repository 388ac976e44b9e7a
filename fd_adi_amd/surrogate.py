"""Learned price surrogate with an ADI fallback.

``MLP`` drives the device network of csrc/mlp.hip (BF16 weights and
activations, FP32 accumulation and master weights, Adam, MFMA GEMMs).
``Ensemble`` averages several independently initialised networks and gates
each query on their spread; ``SurrogatePricer`` answers a batch of queries
with the ensemble mean where the gate passes and the parameters lie inside
the training box, and with the FP64 batched ADI engine elsewhere.

Features, in order: kappa, theta, sigma, rho, v0, r, T, K/S0 (S0 = 100).
"""

import collections
import ctypes

import numpy as np

from . import _lib
from .SizedArrayPtr import SizedArrayPtr
from .batch import HestonBatch
from .utils import sinh_space_batch

FEATURES = ("kappa", "theta", "sigma", "rho", "v0", "r", "T", "K/S0")
BOX_LO = np.array([0.5, 0.01, 0.1, -0.9, 0.01, 0.0, 0.1, 0.7])
BOX_HI = np.array([5.0, 0.16, 1.0, 0.0, 0.16, 0.08, 3.0, 1.3])
N_IN, N_PAD, N_H, N_LAYERS = 8, 128, 512, 4


def nparam():
    n = _lib.c_i64()
    _lib.call("fdadi_mlp_nparam", ctypes.byref(n))
    return n.value


def init_params(seed, n_in=N_IN):
    """He-normal weights for SiLU layers, zero biases (the padded input
    columns beyond the n_in features stay zero and receive zero gradient)."""
    rng = np.random.default_rng(seed)
    parts = []
    for l in range(N_LAYERS):
        fan_in = n_in if l == 0 else N_H
        W = rng.normal(0.0, np.sqrt(2.0 / fan_in), (N_H, N_PAD if l == 0 else N_H))
        if l == 0:
            W[:, n_in:] = 0.0
        parts += [W.ravel(), np.zeros(N_H)]
    parts += [rng.normal(0.0, np.sqrt(1.0 / N_H), N_H), np.zeros(1)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == nparam()
    return p


class DeviceInts(object):
    """int32 device buffer (stored in a double allocation)."""

    def __init__(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        pad = np.zeros(a.size + (a.size % 2), dtype=np.int32)
        pad[:a.size] = a
        self.n = a.size
        self.buf = SizedArrayPtr(pad.view(np.float64))

    @property
    def p(self):
        return self.buf.p

    def at(self, k):
        """Device pointer to element k."""
        base = self.buf.p.value if hasattr(self.buf.p, "value") else int(self.buf.p)
        return ctypes.c_void_p(base + 4 * int(k))

    def to_numpy(self):
        return self.buf.to_numpy().view(np.int32)[:self.n].copy()


class _At(object):
    """A device address with the ``p`` of SizedArrayPtr (a view into a buffer)."""

    __slots__ = ("p",)

    def __init__(self, addr):
        self.p = ctypes.c_void_p(int(addr))


class _IntView(object):
    """int32 view of a device buffer with the ``at`` of DeviceInts."""

    def __init__(self, buf):
        self.buf = buf
        self.p = buf.p

    def at(self, k):
        return ctypes.c_void_p(_addr(self.buf) + 4 * int(k))


def _addr(buf):
    p = buf.p
    return p.value if hasattr(p, "value") else int(p or 0)


def _offset(buf, k, itemsize=8):
    """View of ``buf`` starting at element k."""
    return _At(_addr(buf) + itemsize * int(k))


class Workspace(object):
    """Device scratch buffers kept across calls, grown on demand (sizes in
    doubles)."""

    def __init__(self):
        self._bufs = {}

    def get(self, name, n):
        b = self._bufs.get(name)
        if b is None or b.size < n:
            b = SizedArrayPtr().alloc(max(int(n), 1))
            self._bufs[name] = b
        return b


class Forest(object):
    """The regression trees of a fitted scikit-learn
    ``HistGradientBoostingRegressor`` on the device (csrc/serve.hip): the
    gate's error model evaluated where the queries are.  Prediction is the
    model's baseline plus the reached leaves, tree by tree (scikit-learn's
    order)."""

    def __init__(self, model, nfeat):
        trees = [it[0].nodes for it in model._predictors]
        nodes = np.concatenate(trees)
        if nodes["is_categorical"].any():
            raise NotImplementedError("Forest: categorical splits are not supported")
        start = np.zeros(len(trees) + 1, dtype=np.int64)
        start[1:] = np.cumsum([t.shape[0] for t in trees])
        self._keep = [np.ascontiguousarray(a) for a in (
            start, nodes["feature_idx"].astype(np.int32), nodes["num_threshold"].astype(float),
            nodes["left"].astype(np.int32), nodes["right"].astype(np.int32),
            nodes["is_leaf"].astype(np.uint8), nodes["value"].astype(float))]
        self.base = float(np.ravel(model._baseline_prediction)[0])
        self.nfeat = int(nfeat)
        h = ctypes.c_void_p()
        _lib.call("fdadi_forest_create", len(trees), *[_lib.ptr(a) for a in self._keep[:1]],
                  self.nfeat, *[_lib.ptr(a) for a in self._keep[1:]], self.base,
                  ctypes.byref(h))
        self._h = h

    def __del__(self):
        try:
            if self._h:
                _lib.lib.fdadi_forest_destroy(self._h)
        except Exception:
            pass
        self._h = None

    def predict_device(self, X, n, out):
        _lib.call("fdadi_forest_predict", self._h, X.p, int(n), out.p, _lib.default_stream())

    def predict(self, X_host):
        X_host = np.ascontiguousarray(X_host, dtype=float)
        n = X_host.shape[0]
        X, out = SizedArrayPtr(X_host), SizedArrayPtr().alloc(max(n, 1))
        self.predict_device(X, n, out)
        return out.to_numpy()[:n]


def log_features(P):
    """Monotone per-feature map: the logarithm of the positive scale
    parameters (kappa, theta, sigma, v0, T, K/S0); rho and r unchanged.  It
    spreads the small maturities and variances, where the price moves
    fastest, over more of the network's input range."""
    Q = np.array(P, dtype=float, copy=True)
    for k in (0, 1, 2, 4, 6, 7):
        Q[:, k] = np.log(Q[:, k])
    return Q


FEATURE_MAPS = {"identity": lambda P: np.asarray(P, dtype=float), "log": log_features}


class MLP(object):
    """One network on the device.  Inputs are the 8 features after the
    feature map ``fmap`` (FEATURE_MAPS), normalised with the mapped box
    [lo, hi]; the target with (tmean, tstd)."""

    def __init__(self, lo=BOX_LO, hi=BOX_HI, tmean=0.0, tstd=1.0, seed=0, max_rows=8192,
                 params=None, fmap="identity"):
        self.n_in = len(lo)
        p = (init_params(seed, self.n_in) if params is None
             else np.ascontiguousarray(params, np.float32))
        self.fmap = FEATURE_MAPS[fmap]
        self.fmap_name = fmap
        self.lo = np.ascontiguousarray(self.fmap(np.asarray(lo, float)[None])[0])
        self.hi = np.ascontiguousarray(self.fmap(np.asarray(hi, float)[None])[0])
        self.tmean, self.tstd = float(tmean), float(tstd)
        self.max_rows = int(max_rows)
        h = ctypes.c_void_p()
        _lib.call("fdadi_mlp_create_features", self.n_in, self.max_rows, _lib.ptr(p), p.size,
                  _lib.ptr(self.lo), _lib.ptr(self.hi), self.tmean, self.tstd, ctypes.byref(h))
        self._h = h
        self.stream = _lib.default_stream()

    def __del__(self):
        try:
            if self._h:
                _lib.lib.fdadi_mlp_destroy(self._h)
        except Exception:
            pass
        self._h = None

    def params(self):
        p = np.zeros(nparam(), dtype=np.float32)
        _lib.call("fdadi_mlp_get_params", self._h, _lib.ptr(p), None)
        return p

    def grads(self):
        g = np.zeros(nparam(), dtype=np.float32)
        _lib.call("fdadi_mlp_get_params", self._h, None, _lib.ptr(g))
        return g

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _lib.call("fdadi_mlp_set_params", self._h, _lib.ptr(p), self.stream)

    def grad(self, P, T, idx=None, rows=None):
        """Loss and gradient on device arrays P [n x n_in], T [n] (SizedArrayPtr)."""
        rows = (idx.n if idx is not None else P.size // self.n_in) if rows is None else rows
        loss = ctypes.c_float()
        _lib.call("fdadi_mlp_grad", self._h, P.p, T.p, idx.p if idx is not None else None,
                  int(rows), ctypes.byref(loss), self.stream)
        return loss.value

    def train_step(self, P, T, idx, rows, lr, start=0):
        """One Adam step on rows idx[start:start+rows] (all rows if idx is None)."""
        ip = None if idx is None else (idx.at(start) if start else idx.p)
        _lib.call("fdadi_mlp_train_step", self._h, P.p, T.p, ip, int(rows), ctypes.c_float(lr),
                  self.stream)

    def last_loss(self):
        loss = ctypes.c_float()
        _lib.call("fdadi_mlp_loss", self._h, ctypes.byref(loss))
        return loss.value

    def predict_device(self, P, n, out):
        _lib.call("fdadi_mlp_predict", self._h, P.p, int(n), out.p, self.stream)

    def predict(self, P_host):
        P_host = np.ascontiguousarray(self.fmap(P_host), dtype=float)
        n = P_host.shape[0]
        P = SizedArrayPtr(P_host)
        out = SizedArrayPtr().alloc(n)
        self.predict_device(P, n, out)
        return out.to_numpy()

    def fit(self, P_host, T_host, steps, batch=8192, lr=1e-3, seed=0, log_every=0,
            schedule="constant"):
        """Adam on (P, T) for ``steps`` minibatches drawn without replacement
        per epoch, at rate ``lr`` ("constant") or decayed from ``lr`` to lr/100
        along a half cosine ("cosine"); returns the loss history
        (standardised MSE)."""
        if schedule not in ("constant", "cosine"):
            raise ValueError("schedule must be 'constant' or 'cosine'")
        P = SizedArrayPtr(np.ascontiguousarray(self.fmap(P_host), dtype=float))
        T = SizedArrayPtr(np.ascontiguousarray(T_host, dtype=float))
        n = T_host.shape[0]
        rng = np.random.default_rng(seed)
        rows = min(batch, n)
        per_epoch = max(n // rows, 1)
        # one device permutation per epoch; minibatches are slices of it
        perm, pos, hist = DeviceInts(rng.permutation(n)), 0, []
        for k in range(steps):
            if pos == per_epoch:
                perm, pos = DeviceInts(rng.permutation(n)), 0
            lr_k = lr
            if schedule == "cosine":
                lr_k = lr * (0.01 + 0.99 * 0.5 * (1.0 + np.cos(np.pi * k / max(steps - 1, 1))))
            self.train_step(P, T, perm, rows, lr_k, start=pos * rows)
            pos += 1
            if log_every and (k % log_every == 0 or k == steps - 1):
                hist.append((k, self.last_loss()))
        return hist


def lower_bound(P):
    """European call lower bound max(S0 - K exp(-rT), 0) with S0 = 100: the
    networks learn the price above it (a smaller, smoother target)."""
    P = np.asarray(P, dtype=float)
    return np.maximum(100.0 - 100.0 * P[:, 7] * np.exp(-P[:, 5] * P[:, 6]), 0.0)


def cf_price(P):
    """Semi-analytic Heston call price of feature rows P (S0 = 100), priced on
    the device (fd_adi_amd.pricing): the networks learn the FP64 ADI price's
    difference from it, i.e. the discretisation error of the label mesh."""
    from .pricing import heston_call
    P = np.asarray(P, dtype=float)
    if P.shape[0] == 0:
        return np.zeros(0)
    return heston_call(100.0, 100.0 * P[:, 7], P[:, 5], P[:, 4], P[:, 6], P[:, 0], P[:, 1],
                       P[:, 2], P[:, 3])


# what the networks learn the label's difference from
BASELINES = {"cf": cf_price, "bound": lower_bound}


EnsembleStats = collections.namedtuple("EnsembleStats", "mean spread corr range")
EnsembleStats.__doc__ = """Per-row statistics of an ensemble's predictions (``Ensemble.stats``)."""


def _rows(st, sel):
    """The rows ``sel`` of an ``EnsembleStats``."""
    return EnsembleStats(*[np.asarray(a)[sel] for a in st])


class Ensemble(object):
    """Independently initialised networks; the spread of their predictions
    is the per-query error estimate.  Prices are baseline + network, with the
    baseline the semi-analytic price ("cf") or the call's lower bound
    ("bound")."""

    def __init__(self, members, base="cf"):
        self.members = list(members)
        self.base = base

    def predict(self, P_host):
        st = self.stats(P_host)
        return st.mean, st.spread

    def stats(self, P_host):
        """``EnsembleStats`` of feature rows: the price (baseline + the
        members' mean correction), the members' standard deviation (M - 1),
        their mean correction and their range (max - min)."""
        Y = np.stack([m.predict(P_host) for m in self.members])
        spread = Y.std(axis=0, ddof=1) if len(self.members) > 1 else 0 * Y[0]
        corr = Y.mean(axis=0)
        return EnsembleStats(BASELINES[self.base](P_host) + corr, spread, corr,
                             Y.max(axis=0) - Y.min(axis=0))

    def correction(self, P_host):
        """The members' mean without the baseline (a residual stage)."""
        return np.mean([m.predict(P_host) for m in self.members], axis=0)


VAR_MAX = 1.0   # top of the label variance mesh (above 5x the box's largest variance)


def label_grids(P, nS=256, nv=128):
    """Meshes for feature rows P (vectorised): arrays [n, nS] and [n, nv].
    The spot mesh is K times one unit sinh mesh on [0, 5] clustered at 1
    (density 0.07, i.e. 7 at K = 100) with K on a node; the variance mesh is
    the same power-law mesh v_j = VAR_MAX (j / (nv - 1))^4 for every row.
    Neither depends on S0 or v0, which are not forced onto nodes: the label
    is interpolated there (``label_price``).  Meshes that move with the
    parameters (nodes forced onto S0 and v0, as the engine's own construction
    does) make the label jump by ~1e-3 between neighbouring parameter values;
    these keep it smooth, which is what a regressor needs."""
    P = np.asarray(P, dtype=float)
    K = 100.0 * P[:, 7]
    unit = _unit_spot_mesh(nS)
    v = VAR_MAX * (np.arange(nv) / float(nv - 1)) ** 4
    return K[:, None] * unit[None, :], np.broadcast_to(v, (P.shape[0], nv)).copy()


_UNIT_MESHES = {}


def _unit_spot_mesh(nS):
    if nS not in _UNIT_MESHES:
        _UNIT_MESHES[nS] = sinh_space_batch(np.array([1.0]), 5.0, 0.07, nS)[0]
    return _UNIT_MESHES[nS]


def _lagrange4(x, nodes):
    """Weights [n, 4] of 4-point Lagrange interpolation at x [n] from the
    node rows nodes [n, 4]."""
    w = np.ones(nodes.shape)
    for i in range(4):
        for j in range(4):
            if i != j:
                w[:, i] *= (x - nodes[:, j]) / (nodes[:, i] - nodes[:, j])
    return w


def label_price(V, spots, vars_, S0, v0):
    """Host read-out of grids V [n, nS, nv] at (S0[n], v0[n]) by
    tensor-product cubic Lagrange interpolation on the 4 x 4 nodes around the
    point (exact for bicubic polynomials); meshes spots [n, nS], vars_ [n, nv].
    The labeller does the same on the device (fdadi_hb_read_bicubic)."""
    V = np.asarray(V, dtype=float)
    spots, vars_ = np.asarray(spots, dtype=float), np.asarray(vars_, dtype=float)
    n, nS, nv = V.shape
    S0 = np.broadcast_to(np.asarray(S0, dtype=float), (n,))
    v0 = np.broadcast_to(np.asarray(v0, dtype=float), (n,))
    rows = np.arange(n)
    i = np.clip(np.array([np.searchsorted(spots[k], S0[k]) for k in range(n)]) - 2, 0, nS - 4)
    j = np.clip(np.array([np.searchsorted(vars_[k], v0[k]) for k in range(n)]) - 2, 0, nv - 4)
    di = i[:, None] + np.arange(4)[None, :]
    dj = j[:, None] + np.arange(4)[None, :]
    ws = _lagrange4(S0, spots[rows[:, None], di])
    wv = _lagrange4(v0, vars_[rows[:, None], dj])
    block = V[rows[:, None, None], di[:, :, None], dj[:, None, :]]   # [n, 4, 4]
    return np.einsum("ni,nij,nj->n", ws, block, wv)


def _h2d(dev, host):
    host = np.ascontiguousarray(host, dtype=float)
    _lib.call("fdadi_array_h2d", dev.p, _lib.ptr(host), host.size, _lib.default_stream())


class AdiLabeler(object):
    """FP64 ADI prices of feature rows on the label meshes, computed on the
    device in batches.

    One ``HestonBatch`` per capacity tier (``tiers``: options per batch) and
    its captured graph of ``nt`` steps serve batch after batch: the rows'
    parameters and meshes are loaded with ``HestonBatch.update`` (the graph
    stays valid), the call payoff is written on the device, the graph is
    replayed and the bicubic read-out at (S0 = 100, v0) runs on the device;
    only the prices come back.  A batch shorter than its tier repeats its
    last row (the extra options are discarded).  With ``tune`` the launch
    shapes of each engine are picked by ``HestonBatch.autotune``."""

    def __init__(self, nS=256, nv=128, nt=200, scheme="cs", theta=0.5,
                 tiers=(256, 512, 1024, 2048, 4096), tune=True):
        self.nS, self.nv, self.nt = int(nS), int(nv), int(nt)
        self.scheme, self.theta = scheme, float(theta)
        self.tiers = tuple(sorted(int(t) for t in tiers))
        self.tune = tune
        self._engines = {}

    def _engine(self, cap):
        if cap not in self._engines:
            mid = 0.5 * (BOX_LO + BOX_HI)
            rows = np.repeat(mid[None, :], cap, axis=0)
            hb = HestonBatch(*self._batch_args(rows))
            hb.configure(self.scheme, self.theta, 1.0)
            bufs = [SizedArrayPtr().alloc(cap) for _ in range(4)]   # K, S0, v0, out
            _h2d(bufs[0], 100.0 * rows[:, 7])
            hb.call_payoff(bufs[0])
            if self.tune:
                hb.autotune(steps=1)
            hb.capture(self.nt)
            self._engines[cap] = (hb, bufs)
        return self._engines[cap]

    def warm(self):
        """Build every tier's engine, launch shapes and graph now."""
        for t in self.tiers:
            self._engine(t)
        return self

    def _batch_args(self, rows):
        spots, vars_ = label_grids(rows, self.nS, self.nv)
        params = dict(kappa=rows[:, 0], theta=rows[:, 1], sigma=rows[:, 2], rho=rows[:, 3],
                      r=rows[:, 5])
        return params, spots, vars_, rows[:, 6] / self.nt

    def _tier(self, m):
        for t in self.tiers:
            if m <= t:
                return t
        return self.tiers[-1]

    def _batches(self, m):
        """Batch sizes (tiers) for m rows, in order: full batches of the
        largest tier, then the rest as the smallest tier that holds it when
        that pads by at most an eighth of the tier, else the largest tier
        that fits in it, and so on (padding rows cost as much as real ones)."""
        caps, rem, big = [], int(m), self.tiers[-1]
        while rem >= big:
            caps.append(big)
            rem -= big
        while rem > 0:
            t = self._tier(rem)
            lower = [x for x in self.tiers if x <= rem]
            if t - rem <= t // 8 or not lower:
                caps.append(t)
                rem = 0
            else:
                caps.append(lower[-1])
                rem -= lower[-1]
        return caps

    def price(self, P, timed=False):
        """FP64 ADI price of every feature row of P (S0 = 100).  With
        ``timed`` one batch (the first of the largest tier, else the last)
        is stepped stage by stage with a HIP event pair around each launch
        instead of the graph (the same kernels): ``last_stage_ms`` then holds
        the device milliseconds per stage slot summed over the nt steps,
        ``last_batch`` that batch's option count."""
        P = np.ascontiguousarray(P, dtype=float)
        n = P.shape[0]
        if n == 0:
            return np.zeros(0)
        Pd, out = SizedArrayPtr(P), SizedArrayPtr().alloc(n)
        self.price_rows_device(Pd, None, n, out, timed=timed)
        return out.to_numpy()[:n]

    def price_host_batches(self, P, timed=False):
        """``price`` with each batch's parameters and meshes built on the host
        and loaded by ``HestonBatch.update`` (the engine's host path; kept for
        tests of that path)."""
        P = np.asarray(P, dtype=float)
        out = np.empty(P.shape[0])
        c0 = 0
        pending = bool(timed)
        while c0 < P.shape[0]:
            cap = self._tier(P.shape[0] - c0)
            rows = P[c0:c0 + cap]
            m = rows.shape[0]
            if m < cap:
                rows = np.vstack([rows, np.repeat(rows[-1:], cap - m, axis=0)])
            hb, (K, S0, v0, res) = self._engine(cap)
            hb.update(*self._batch_args(rows))
            _h2d(K, 100.0 * rows[:, 7])
            _h2d(S0, np.full(cap, 100.0))
            _h2d(v0, rows[:, 4])
            hb.call_payoff(K)
            if pending and (cap == self.tiers[-1] or c0 + m >= P.shape[0]):
                self.last_stage_ms = hb.step_timed(self.nt)
                self.last_batch = cap
                pending = False
            else:
                hb.replay()
            hb.read_bicubic(S0, v0, res)
            out[c0:c0 + m] = res.to_numpy()[:m]
            c0 += m
        return out

    def price_rows_device(self, P, idx, m, out, timed=False):
        """FP64 ADI prices of rows idx[0..m) (a ``DeviceInts``; None: rows
        0..m) of the device feature rows P, scattered into out[idx[i]] (out[i]
        without idx); nothing but the launches leaves the host.  ``timed`` as
        in ``price`` (the first batch of the largest tier, else the last)."""
        if m <= 0:
            return
        if not hasattr(self, "_unit"):
            self._unit = SizedArrayPtr(np.ascontiguousarray(_unit_spot_mesh(self.nS)))
        s = _lib.default_stream()
        c0, pending = 0, bool(timed)
        for cap in self._batches(m):
            mm = min(cap, m - c0)
            hb, (K, S0, v0, res) = self._engine(cap)
            ws = self._dev_ws(cap)
            ip = None if idx is None else idx.at(c0)
            Pc = P if idx is not None else _offset(P, 8 * c0)
            _lib.call("fdadi_serve_label_inputs", Pc.p, ip, int(mm), int(cap), self.nt,
                      self._unit.p, self.nS, ws["params"].p, ws["spots"].p, K.p, S0.p, v0.p, s)
            _lib.call("fdadi_hb_update_dev", hb._h, ws["params"].p, ws["spots"].p, None, s)
            hb.call_payoff(K)
            if pending and (cap == self.tiers[-1] or c0 + mm >= m):
                self.last_stage_ms = hb.step_timed(self.nt)
                self.last_batch = cap
                pending = False
            else:
                hb.replay()
            hb.read_bicubic(S0, v0, res)
            if idx is None:
                _lib.call("fdadi_array_copy", _offset(out, c0).p, res.p, int(mm), s)
            else:
                _lib.call("fdadi_serve_scatter", res.p, ip, int(mm), out.p, s)
            c0 += mm

    def _dev_ws(self, cap):
        if not hasattr(self, "_ws"):
            self._ws = {}
        if cap not in self._ws:
            self._ws[cap] = {"params": SizedArrayPtr().alloc(6 * cap),
                             "spots": SizedArrayPtr().alloc(self.nS * cap)}
        return self._ws[cap]

    def grids(self, P):
        """Value grids [n, nS, nv] of feature rows P (one tier-sized batch at
        most; for tests)."""
        P = np.asarray(P, dtype=float)
        cap = self._tier(P.shape[0])
        if P.shape[0] > cap:
            raise ValueError("grids: at most %d rows" % cap)
        m = P.shape[0]
        rows = np.vstack([P, np.repeat(P[-1:], cap - m, axis=0)]) if m < cap else P
        hb, (K, S0, v0, res) = self._engine(cap)
        hb.update(*self._batch_args(rows))
        _h2d(K, 100.0 * rows[:, 7])
        hb.call_payoff(K)
        hb.replay()
        return np.stack(hb.get_state()[:m])


_LABELERS = {}


def labeler(nS=256, nv=128, nt=200, scheme="cs"):
    """The process's ``AdiLabeler`` for one grid shape (created on first use)."""
    key = (int(nS), int(nv), int(nt), scheme)
    if key not in _LABELERS:
        _LABELERS[key] = AdiLabeler(nS, nv, nt, scheme)
    return _LABELERS[key]


def adi_prices(P, nS=256, nv=128, nt=200, scheme="cs", timed=False):
    """FP64 ADI prices (batched engine on the device) for feature rows P on
    the label meshes, read at (S0 = 100, v0)."""
    return labeler(nS, nv, nt, scheme).price(P, timed=timed)


def within_call_bounds(P, y, slack=1.0):
    """Rows whose price lies within the European call bounds
    max(S0 - K exp(-rT), 0) <= C <= S0 (S0 = 100), widened by ``slack``.

    The ADI schemes on a coarse label mesh are unstable for some
    advection-dominated variance dynamics (large kappa*theta, small sigma);
    such labels leave these bounds by orders of magnitude and are dropped
    from the training set."""
    P = np.asarray(P, dtype=float)
    y = np.asarray(y, dtype=float)
    lower = np.maximum(100.0 - 100.0 * P[:, 7] * np.exp(-P[:, 5] * P[:, 6]), 0.0)
    return np.isfinite(y) & (y >= lower - slack) & (y <= 100.0 + slack)


def draw_params(n, seed, outside_frac=0.0, widen=0.25):
    """``n`` feature rows drawn uniformly over the training box; a fraction
    ``outside_frac`` of them instead lies outside it: drawn from the box
    widened by ``widen`` of each range on both sides until at least one
    feature is outside, then clipped to physically valid values (positive
    variances, vol-of-vol, mean reversion and maturity; |rho| < 1)."""
    rng = np.random.default_rng(seed)
    P = rng.uniform(BOX_LO, BOX_HI, (n, N_IN))
    n_out = int(round(outside_frac * n))
    if n_out:
        w = BOX_HI - BOX_LO
        lo, hi = BOX_LO - widen * w, BOX_HI + widen * w
        rows = []
        while len(rows) < n_out:
            x = rng.uniform(lo, hi, (2 * n_out, N_IN))
            x = x[np.any((x < BOX_LO) | (x > BOX_HI), axis=1)]
            rows.extend(list(x))
        Q = np.array(rows[:n_out])
        Q[:, 0] = np.maximum(Q[:, 0], 0.1)     # kappa
        Q[:, 1] = np.maximum(Q[:, 1], 0.005)   # theta
        Q[:, 2] = np.maximum(Q[:, 2], 0.05)    # sigma
        Q[:, 3] = np.clip(Q[:, 3], -0.99, 0.99)
        Q[:, 4] = np.maximum(Q[:, 4], 0.005)   # v0
        Q[:, 6] = np.maximum(Q[:, 6], 0.05)    # T
        Q[:, 7] = np.maximum(Q[:, 7], 0.3)     # K / S0
        P[rng.choice(n, n_out, replace=False)] = Q
    return P


def feller_ratio(P):
    """2 kappa theta / sigma^2 of feature rows P.  Below ~1.5 the variance
    process crowds v = 0, where the label mesh resolves it poorly: there the
    FP64 ADI labels differ from the exact price by 1e-2 and more, against
    ~1e-3 elsewhere."""
    P = np.asarray(P, dtype=float)
    return 2.0 * P[:, 0] * P[:, 1] / P[:, 2] ** 2


class Stack(object):
    """Ensembles added up: the first predicts the price (its baseline plus
    its mean), each later one a correction of the residual left by those
    before it (trained with ``offset``); the spread is the first's."""

    def __init__(self, stages):
        self.stages = list(stages)
        self.members = self.stages[0].members

    def predict(self, P_host):
        price, spread = self.stages[0].predict(P_host)
        for st in self.stages[1:]:
            price = price + st.correction(P_host)
        return price, spread


def train_ensemble(P, y, members=4, steps=2000, batch=8192, lr=1e-3, seed=0, log_every=0,
                   base="cf", max_resid=None, schedule="constant", subsample=1.0,
                   fmap="identity", offset=None):
    """Independently initialised networks trained on the same labels (each on
    its own minibatch order) to predict y - BASELINES[base](P); with
    ``max_resid`` the rows whose target exceeds it in magnitude are left out
    (labels far from the baseline come from regimes the label mesh does not
    resolve; the gate sends those queries to ADI).  With ``subsample`` < 1
    each member trains on its own random subset of that fraction of the rows
    (bagging), so the members' spread also reflects how well the data pin
    the function.  Returns the Ensemble and the loss histories."""
    P = np.asarray(P, dtype=float)
    t = np.asarray(y, dtype=float) - BASELINES[base](P)
    if offset is not None:
        t = t - np.asarray(offset, dtype=float)
    if max_resid is not None:
        keep = np.abs(t) <= max_resid
        P, t = P[keep], t[keep]
    mu, sd = float(np.mean(t)), float(np.std(t))
    nets, hists = [], []
    if not 0.0 < subsample <= 1.0:
        raise ValueError("subsample must be in (0, 1]")
    for k in range(members):
        m = MLP(BOX_LO, BOX_HI, mu, sd, seed=seed + 101 * k, max_rows=batch, fmap=fmap)
        Pk, tk = P, t
        if subsample < 1.0:
            pick = np.random.default_rng(seed + 13 * k).permutation(len(t))
            pick = np.sort(pick[:max(int(subsample * len(t)), 1)])
            Pk, tk = P[pick], t[pick]
        hists.append(m.fit(Pk, tk, steps, batch=batch, lr=lr, seed=seed + 7 * k,
                           log_every=log_every, schedule=schedule))
        nets.append(m)
    return Ensemble(nets, base), hists


def calibrate_threshold(score, err, atol=1e-4, max_fail=0.0):
    """Largest threshold tau on the per-query error estimate ``score`` such
    that, among calibration queries with score <= tau, the fraction whose
    actual error ``err`` exceeds ``atol`` is at most ``max_fail`` (0: none).
    Ties are respected: tau always sits at a score value below which every
    query with that value is admitted.  Returns (tau, accepted fraction of
    the calibration set); tau = -inf (admit nothing) when no threshold
    qualifies."""
    score = np.asarray(score, dtype=float)
    err = np.asarray(err, dtype=float)
    if score.size == 0:
        return -np.inf, 0.0
    order = np.argsort(score, kind="stable")
    s, bad = score[order], np.cumsum(err[order] > atol)
    n = np.arange(1, s.size + 1)
    ok = bad <= max_fail * n
    # a cut after position k is valid only where the next score is larger
    ok[:-1] &= s[1:] > s[:-1]
    ok &= np.isfinite(s)
    idx = np.nonzero(ok)[0]
    if idx.size == 0:
        return -np.inf, 0.0
    k = int(idx[-1])
    return float(s[k]), (k + 1) / float(s.size)


def calibration_spread(score, err, atol=1e-4, max_fail=0.0, n=200, seed=0,
                       q=(0.05, 0.5, 0.95)):
    """Quantiles q of the accepted fraction that ``calibrate_threshold``
    gives over n bootstrap resamples (rows drawn with replacement) of the
    calibration rows: how far the admission could move with another draw
    of those rows.  A zero-failure threshold is set by the lowest-scored
    failing rows, so this spread is wide where few failures lie near it."""
    score = np.asarray(score, dtype=float)
    err = np.asarray(err, dtype=float)
    if score.size == 0:
        return {"%g" % a: 0.0 for a in q}
    rng = np.random.default_rng(seed)
    acc = np.empty(n)
    for b in range(n):
        i = rng.integers(0, score.size, score.size)
        acc[b] = calibrate_threshold(score[i], err[i], atol, max_fail)[1]
    return {"%g" % a: float(v) for a, v in zip(q, np.quantile(acc, q))}


def calibrate_gate(ensemble, P, y, atol=1e-4, feller_min=0.0, max_fail=0.0):
    """Spread threshold of an ensemble (at least two members) calibrated on
    held-out rows P with FP64 ADI labels y (``calibrate_threshold``); rows
    with Feller ratio below ``feller_min`` are never admitted.  Returns (tau,
    accepted fraction of all the rows)."""
    if len(ensemble.members) < 2:
        raise ValueError("a spread gate needs an ensemble of at least two networks")
    mean, spread = ensemble.predict(P)
    err = np.abs(mean - np.asarray(y, dtype=float))
    region = feller_ratio(P) >= feller_min
    tau, frac = calibrate_threshold(spread[region], err[region], atol, max_fail)
    return tau, frac * region.mean()


class ErrorModel(object):
    """Per-query estimate of the surrogate's error, fitted on held-out FP64
    ADI labels: a regression of log10 |surrogate - ADI| on the 8 features,
    the log Feller ratio and four statistics of the ensemble's corrections
    (log spread, mean, log |mean|, log range).  The members' spread alone
    cannot see the error they share; this model learns where it lies (low
    Feller ratio, large vol-of-vol, long maturities near the money, large
    corrections).

    ``kind`` "trees" (default): gradient-boosted regression trees
    (scikit-learn's histogram GBM; evaluated on the device by ``Forest`` when
    serving); "mlp": a network of the library's own MFMA kernels on the same
    features.  The trees do far better than the network at the same
    calibration (the regions where the error jumps have sharp edges in the
    parameters, which trees follow).  On the held-out rows of 2^20 labels and
    8 members, adding the mean correction, its magnitude and the range to the
    spread raised the admission at no calibration failure from 70% to 75%
    (mean of the two halves of the calibration rows, each calibrating the
    other)."""

    FLOOR = 1e-9
    NFEAT = 13

    def __init__(self, kind="trees", models=4):
        if kind not in ("trees", "mlp"):
            raise ValueError("ErrorModel kind must be 'trees' or 'mlp'")
        self.kind = kind
        self.n_models = int(models) if kind == "trees" else 1
        if self.n_models < 1:
            raise ValueError("ErrorModel needs at least one model")
        self.models = []

    @staticmethod
    def features(P, st):
        """Feature rows [n, 13] of rows P and their ``EnsembleStats`` st (the
        columns fdadi_serve_combine writes on the device)."""
        P = np.asarray(P, dtype=float)
        corr = np.asarray(st.corr, dtype=float)
        return np.column_stack([log_features(P), np.log(feller_ratio(P)),
                                np.log10(np.asarray(st.spread, dtype=float) + 1e-12), corr,
                                np.log10(np.abs(corr) + 1e-9),
                                np.log10(np.asarray(st.range, dtype=float) + 1e-12)])

    def fit(self, P, st, err, steps=4000, batch=8192, lr=1e-3, seed=0):
        """Trees: ``n_models`` boosted-tree regressors that differ only in
        their random state (the split of their early-stopping rows); the
        score of a row is the largest of their estimates.  One model's
        zero-failure admission depends strongly on that state (on the 2^20
        label set, 23% to 72% of the calibration rows over five states); the
        largest of four held 71-77% over three sets of states."""
        X = self.features(P, st)
        t = np.log10(np.asarray(err, dtype=float) + self.FLOOR)
        self.models = []
        if self.kind == "trees":
            from sklearn.ensemble import HistGradientBoostingRegressor
            for m in range(self.n_models):
                model = HistGradientBoostingRegressor(max_iter=300, learning_rate=0.1,
                                                      max_leaf_nodes=63, random_state=seed + m)
                self.models.append(model.fit(X, t))
            return []
        lo, hi = X.min(axis=0), X.max(axis=0)
        pad = 0.05 * (hi - lo) + 1e-6
        net = MLP(lo - pad, hi + pad, float(t.mean()), float(t.std()), seed=seed, max_rows=batch)
        self.models = [net]
        return net.fit(X, t, steps, batch=batch, lr=lr, seed=seed, schedule="cosine")

    def __call__(self, P, st):
        X = self.features(P, st)
        return np.max(np.stack([m.predict(X) for m in self.models]), axis=0)

    def forest(self):
        """The fitted trees on the device, one ``Forest`` per model (kind
        "trees" only)."""
        if self.kind != "trees":
            raise NotImplementedError("ErrorModel.forest: the trees model only")
        return [Forest(m, self.NFEAT) for m in self.models]


def fit_surrogate(P, y, members=4, steps=16000, batch=8192, fmap="identity", atol=1e-4,
                  max_fail=0.0, holdout=0.25, calib=1.0 / 3.0, error_steps=4000, seed=0,
                  feller_min=0.0, error_kind="trees", calib_margin=0.9):
    """The gated surrogate from FP64 ADI labels y of feature rows P.

    The rows split into a training part (1 - holdout) and a held-out part;
    the ensemble trains on the training rows whose labels are stable (within
    the call bounds and within 0.05 of the semi-analytic price); the error
    model trains on the first (1 - calib) of the held-out rows, on the
    ensemble's actual errors there; the gate threshold is calibrated on the
    rest of the held-out rows (``calibrate_threshold``: at most ``max_fail``
    of the admitted calibration rows beyond ``calib_margin * atol``; the
    margin keeps the threshold's generalisation to new queries within
    ``atol``: on 2^20 labels, calibrating at 1e-4 let 1 to 2 of ~65,000
    admitted fresh queries reach 1.4e-4, at 0.9e-4 none exceeded 1e-4).
    Returns the SurrogatePricer and a dict of diagnostics."""
    P = np.asarray(P, dtype=float)
    y = np.asarray(y, dtype=float)
    n = len(y)
    base = cf_price(P)
    stable = within_call_bounds(P, y) & (np.abs(y - base) < 0.05)
    ntr = int(round((1.0 - holdout) * n))
    tr = np.arange(ntr)[stable[:ntr]]
    ens, hist = train_ensemble(P[tr], y[tr], members=members, steps=steps, batch=batch,
                               schedule="cosine", fmap=fmap, seed=seed)
    H = np.arange(ntr, n)
    st = ens.stats(P[H])
    err = np.abs(st.mean - y[H])
    nfit = int(round((1.0 - calib) * len(H)))
    em = ErrorModel(error_kind)
    em.fit(P[H[:nfit]], _rows(st, slice(0, nfit)), err[:nfit], steps=error_steps, seed=seed + 1)
    Hc = H[nfit:]
    score = em(P[Hc], _rows(st, slice(nfit, None)))
    region = feller_ratio(P[Hc]) >= feller_min
    catol = calib_margin * atol
    tau, frac = calibrate_threshold(score[region], err[nfit:][region], catol, max_fail)
    info = dict(error_model=error_kind, calib_atol=catol, gate_calib_accept_at_budget={
                    "%g" % b: float(calibrate_threshold(score[region], err[nfit:][region], catol,
                                                        b)[1] * region.mean())
                    for b in (1e-4, 1e-3, 1e-2)},
                train_rows=int(len(tr)), heldout_rows=int(len(H)), error_fit_rows=int(nfit),
                calib_rows=int(len(Hc)), labels_unstable=int((~stable).sum()),
                heldout_frac_within={"%g" % a: float(np.mean(err <= a))
                                     for a in (1e-5, 1e-4, 1e-3)},
                calib_accept=float(frac * region.mean()), tau=float(tau),
                calib_accept_bootstrap={
                    k: v * float(region.mean()) for k, v in calibration_spread(
                        score[region], err[nfit:][region], catol, max_fail).items()},
                train_final_loss=[float(h[-1][1]) for h in hist if h])
    return SurrogatePricer(ens, tau, feller_min=feller_min, error_model=em), info


class SurrogatePricer(object):
    """Ensemble surrogate with a per-query gate and the FP64 ADI fallback.

    A query is answered by the surrogate (semi-analytic price + ensemble
    mean) when it lies inside the training box, its Feller ratio is at least
    ``feller_min`` and its error estimate is at most ``tau``; every other
    query goes to ``adi`` (default: ``adi_prices`` on the label meshes).  The
    error estimate is the ensemble spread, or ``error_model(P, stats)`` when
    given (an ``ErrorModel``, or any callable of the rows and their
    ``EnsembleStats`` returning one estimate per row)."""

    def __init__(self, ensemble, tau, lo=BOX_LO, hi=BOX_HI, feller_min=0.0, error_model=None,
                 adi=None):
        self.last_fallback = 0
        if error_model is None and len(ensemble.members) < 2:
            raise ValueError("a spread gate needs an ensemble of at least two networks")
        self.ensemble, self.tau = ensemble, float(tau)
        self.lo, self.hi = np.asarray(lo, float), np.asarray(hi, float)
        self.feller_min = float(feller_min)
        self.error_model = error_model
        self.adi = adi if adi is not None else adi_prices

    def score(self, P, st):
        """Error estimate of rows P from their ``EnsembleStats`` st."""
        return st.spread if self.error_model is None else self.error_model(P, st)

    def gate(self, P, score):
        inside = np.all((P >= self.lo) & (P <= self.hi), axis=1)
        return inside & (feller_ratio(P) >= self.feller_min) & (score <= self.tau)

    def price(self, P, **adi_kw):
        """Prices and the admitted flags of host feature rows P.  With the
        trees error model and the default fallback the batch goes through
        ``price_device``; otherwise the gate runs on the host."""
        P = np.asarray(P, dtype=float)
        if self.device_ready():
            n = P.shape[0]
            if n == 0:
                return np.zeros(0), np.zeros(0, dtype=bool)
            Pd = SizedArrayPtr(np.ascontiguousarray(P))
            out, ok = SizedArrayPtr().alloc(n), DeviceInts(np.zeros(n, dtype=np.int32))
            self.price_device(Pd, n, out, ok, **adi_kw)
            return out.to_numpy()[:n], ok.to_numpy()[:n].astype(bool)
        st = self.ensemble.stats(P)
        ok = self.gate(P, self.score(P, st))
        out = st.mean.copy()
        self.last_fallback = int((~ok).sum())
        if (~ok).any():
            out[~ok] = self.adi(P[~ok], **adi_kw)
        return out, ok

    def device_ready(self):
        """Whether the whole batch can be served on the device: the trees
        error model, identity-mapped members on the semi-analytic baseline,
        and the default ADI fallback."""
        em = self.error_model
        return (isinstance(em, ErrorModel) and em.kind == "trees" and self.adi is adi_prices and
                self.ensemble.base == "cf" and
                all(m.fmap_name == "identity" for m in self.ensemble.members))

    def price_device(self, P, n, out, ok, timed=False, labeler_kw=None):
        """Answer n feature rows P (device, 8 doubles each) into out (device,
        n prices) and ok (a ``DeviceInts`` of n admitted flags): the members'
        corrections, the semi-analytic baseline, the gate's features and
        trees, the gate and the FP64 ADI fallback of the rejected rows all run
        on the device; only the count of rejected rows comes back.
        ``timed``: the fallback's first batch is stepped stage by stage with
        HIP events (``AdiLabeler.price``)."""
        if not self.device_ready():
            raise NotImplementedError("price_device: needs the trees error model, "
                                      "identity-mapped members on the cf baseline and the "
                                      "default fallback")
        if not hasattr(self, "_ws"):
            self._ws = Workspace()
            self._forest = self.error_model.forest()
        ws, M, s = self._ws, len(self.ensemble.members), _lib.default_stream()
        F = len(self._forest)
        Y, base = ws.get("Y", M * n), ws.get("base", n)
        rows, X = ws.get("rows", 9 * n), ws.get("X", ErrorModel.NFEAT * n)
        score, scores = ws.get("score", n), ws.get("scores", F * n)
        inside, idx = ws.get("inside", (n + 1) // 2), ws.get("idx", (n + 1) // 2)
        work = ws.get("work", (n + 255) // 256 + 2)
        for k, m in enumerate(self.ensemble.members):
            m.predict_device(P, n, _offset(Y, k * n))
        _lib.call("fdadi_serve_cf_rows", P.p, int(n), rows.p, s)
        _lib.call("fdadi_heston_cf_call", rows.p, int(n), base.p, s)
        lo, hi = np.ascontiguousarray(self.lo), np.ascontiguousarray(self.hi)
        _lib.call("fdadi_serve_combine", P.p, int(n), Y.p, M, base.p, _lib.ptr(lo), _lib.ptr(hi),
                  self.feller_min, out.p, X.p, inside.p, s)
        # the error models' estimates, and the largest of them per row
        for f, forest in enumerate(self._forest):
            forest.predict_device(X, n, _offset(scores, f * n))
        _lib.call("fdadi_serve_max_rows", scores.p, F, int(n), score.p, s)
        nrej = _lib.c_i32()
        _lib.call("fdadi_serve_gate", score.p, inside.p, int(n), self.tau, ok.p, idx.p, work.p,
                  ctypes.byref(nrej), s)
        self.last_fallback = int(nrej.value)
        lab = labeler(**(labeler_kw or {}))
        lab.price_rows_device(P, _IntView(idx), self.last_fallback, out, timed=timed)
