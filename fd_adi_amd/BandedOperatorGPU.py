"""Device banded operators -- drop-in for FiniteDifference/BandedOperatorGPU.pyx.

``BandedOperator`` wraps exactly one native operator: a tridiagonal operator
(``fdadi_tri``, replacing _TriBandedOperator) or the 9-point cross-derivative
operator (``fdadi_cross``, replacing _CSRBandedOperator), keeping the
reference's method names, argument meaning and error behaviour
(BandedOperatorGPU.pyx:27-641).

Host-side operators (for ``emigrate``/``immigrate``) are duck-typed: any object
with the CPU BandedOperator attributes (``D`` or ``data``/``offsets`` in scipy
DIA layout, ``R``, ``dirichlet``, ``blocks``, ``axis``, fold statuses and
factors) can be moved to the device; ``immigrate`` returns a
:class:`HostBandedOperator` with those same attributes.
"""

import ctypes

import numpy as np
import scipy.sparse

from . import _lib
from .SizedArrayPtr import SizedArrayPtr, _Buffer

FOLDED = "FOLDED"
CAN_FOLD = "CAN_FOLD"
CANNOT_FOLD = "CANNOT_FOLD"
_STATUS_I = {CANNOT_FOLD: 0, CAN_FOLD: 1, FOLDED: 2}
_STATUS_S = {v: k for k, v in _STATUS_I.items()}

_scratch = {}


def _scratch_buffer(n):
    """Per-size device scratch reused by in-place apply (the reference's
    SizedArray tempspace, VecArray.h:237, without one per array)."""
    b = _scratch.get(n)
    if b is None or not b.alive:
        b = _Buffer(n)
        _scratch[n] = b
    return b


class HostBandedOperator(object):
    """Host copy of an operator in the CPU BandedOperator layout
    (scipy DIA: data[k, j] = A[j - offsets[k], j], offsets descending)."""

    def __init__(self, data, offsets, R, axis=0, blocks=1, derivative=1, order=2,
                 deltas=None):
        self.data = np.asarray(data, dtype=float)
        self.offsets = tuple(int(o) for o in offsets)
        n = self.data.shape[1]
        self.shape = (n, n)
        self.R = None if R is None else np.asarray(R, dtype=float)
        self.axis = axis
        self.blocks = blocks
        self.derivative = derivative
        self.order = order
        self.deltas = np.array([np.nan]) if deltas is None else deltas
        self.dirichlet = [None, None]
        self.top_factors = None
        self.bottom_factors = None
        self.top_fold_status = CANNOT_FOLD
        self.bottom_fold_status = CANNOT_FOLD
        self.is_mixed_derivative = False
        self.solve_banded_offsets = (abs(min(self.offsets)), abs(max(self.offsets)))

    @property
    def D(self):
        return scipy.sparse.dia_matrix((self.data, self.offsets), shape=self.shape)


def _dia_of(other):
    """(data, offsets) of a host operator in scipy DIA layout."""
    if hasattr(other, "data") and hasattr(other, "offsets") and not scipy.sparse.issparse(other):
        return np.asarray(other.data, dtype=float), tuple(int(o) for o in other.offsets)
    D = scipy.sparse.dia_matrix(other.D)
    return np.asarray(D.data, dtype=float), tuple(int(o) for o in D.offsets)


def _row_band(data, offsets, o, n):
    """Row-indexed band: out[r] = A[r, r + o] from scipy DIA data."""
    out = np.zeros(n)
    if o not in offsets:
        return out
    d = data[offsets.index(o)]
    r = np.arange(n)
    c = r + o
    ok = (c >= 0) & (c < n)
    out[r[ok]] = d[c[ok]]
    return out


def _dia_from_rows(bands, offsets, n):
    """scipy DIA data from row-indexed bands (inverse of _row_band)."""
    data = np.zeros((len(offsets), n))
    r = np.arange(n)
    for k, o in enumerate(offsets):
        c = r + o
        ok = (c >= 0) & (c < n)
        data[k, c[ok]] = bands[k][r[ok]]
    return data


class BandedOperator(object):

    def __init__(self, other=None, tag=""):
        self.attrs = ('deltas', 'derivative', 'is_mixed_derivative', 'order', 'axis',
                      'blocks', 'top_fold_status', 'bottom_fold_status')
        self._tri = None
        self._csr = None
        self.deltas = None
        self.derivative = 1
        self.is_mixed_derivative = False
        self.order = 2
        self.axis = 0
        self.blocks = 1
        self.tag = tag
        if other is not None:
            self.emigrate(other, tag)

    # -- native handle plumbing -----------------------------------------------------
    @property
    def thisptr_tri(self):
        return self._tri

    @property
    def thisptr_csr(self):
        return self._csr

    def _tri_info(self):
        vals = [_lib.c_i64(), _lib.c_i64(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(),
                ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), _lib.c_i64()]
        _lib.call("fdadi_tri_info", self._tri, *[ctypes.byref(v) for v in vals])
        return [v.value for v in vals]

    @property
    def top_fold_status(self):
        if self._tri is None:
            return CANNOT_FOLD
        return _STATUS_S[self._tri_info()[5]]

    @top_fold_status.setter
    def top_fold_status(self, val):
        pass  # tracked natively; setting from Python is a no-op

    @property
    def bottom_fold_status(self):
        if self._tri is None:
            return CANNOT_FOLD
        return _STATUS_S[self._tri_info()[6]]

    @bottom_fold_status.setter
    def bottom_fold_status(self, val):
        pass

    @property
    def has_high_dirichlet(self):
        return False if self.is_mixed_derivative else bool(self._tri_info()[4])

    @has_high_dirichlet.setter
    def has_high_dirichlet(self, val):
        if self.is_mixed_derivative:
            raise NotImplementedError("Mixed derivatives cannot have dirichlet boundaries.")
        _lib.call("fdadi_tri_set_flags", self._tri, -1, int(bool(val)), -1)

    @property
    def has_low_dirichlet(self):
        return False if self.is_mixed_derivative else bool(self._tri_info()[3])

    @has_low_dirichlet.setter
    def has_low_dirichlet(self, val):
        if self.is_mixed_derivative:
            raise NotImplementedError("Mixed derivatives cannot have dirichlet boundaries.")
        _lib.call("fdadi_tri_set_flags", self._tri, int(bool(val)), -1, -1)

    @property
    def has_residual(self):
        return True if self.is_mixed_derivative else bool(self._tri_info()[7])

    @property
    def operator_rows(self):
        if self.is_mixed_derivative:
            n0, n1 = _lib.c_i64(), _lib.c_i64()
            _lib.call("fdadi_cross_info", self._csr, ctypes.byref(n0), ctypes.byref(n1))
            return n0.value * n1.value
        return self._tri_info()[0]

    @property
    def shape(self):
        return (self.operator_rows, self.operator_rows)

    def __del__(self):
        try:
            if self._csr is not None:
                _lib.lib.fdadi_cross_destroy(self._csr)
            elif self._tri is not None:
                _lib.lib.fdadi_tri_destroy(self._tri)
        except Exception:
            pass
        self._csr = self._tri = None

    def copy_meta_data(self, other, **kwargs):
        for attr in self.attrs:
            if attr in ('top_fold_status', 'bottom_fold_status'):
                continue
            val = kwargs[attr] if attr in kwargs else getattr(other, attr, None)
            setattr(self, attr, val)

    def __richcmp__(self, other, op):
        raise NotImplementedError

    def __eq__(self, other):
        raise NotImplementedError("We can't and won't do comparisons of operators "
                                  "living on the GPU")

    __hash__ = object.__hash__

    # -- copy / migration --------------------------------------------------------------
    def copy(self):
        B = BandedOperator()
        h = ctypes.c_void_p()
        if self.is_mixed_derivative:
            _lib.call("fdadi_cross_copy", self._csr, _lib.default_stream(), ctypes.byref(h))
            B._csr = h
        else:
            _lib.call("fdadi_tri_copy", self._tri, _lib.default_stream(), ctypes.byref(h))
            B._tri = h
        B.copy_meta_data(self)
        return B

    def emigrate(self, other, tag=""):
        """Become a device copy of the host operator ``other``
        (BandedOperatorGPU.pyx:157-314)."""
        self.copy_meta_data(other)
        self.is_mixed_derivative = bool(getattr(other, "is_mixed_derivative", False))
        if self.is_mixed_derivative:
            return self._emigrate_csr(other)
        return self._emigrate_tri(other)

    def _emigrate_tri(self, other):
        data, offsets = _dia_of(other)
        n = data.shape[1]
        blocks = int(getattr(other, "blocks", 1))
        bl = n // blocks
        extra = set(offsets) - {1, 0, -1, 2, -2}
        if extra:
            raise ValueError("operator has bands %s beyond +-2" % sorted(extra))
        sup = _row_band(data, offsets, 1, n)
        mid = _row_band(data, offsets, 0, n)
        sub = _row_band(data, offsets, -1, n)
        starts = np.arange(blocks) * bl
        top_status = getattr(other, "top_fold_status", CANNOT_FOLD)
        bottom_status = getattr(other, "bottom_fold_status", CANNOT_FOLD)
        tops = np.zeros(blocks)
        bots = np.zeros(blocks)
        # CAN_FOLD: the +-2 band entries are carried in the factor arrays.
        if 2 in offsets:
            tops = _row_band(data, offsets, 2, n)[starts]
            top_status = CAN_FOLD
        elif top_status == FOLDED:
            tops = np.atleast_1d(np.asarray(other.top_factors, dtype=float))
        if -2 in offsets:
            bots = _row_band(data, offsets, -2, n)[starts + bl - 1]
            bottom_status = CAN_FOLD
        elif bottom_status == FOLDED:
            bots = np.atleast_1d(np.asarray(other.bottom_factors, dtype=float))
        R = getattr(other, "R", None)
        has_R = R is not None
        R = np.zeros(n) if R is None else np.ascontiguousarray(R, dtype=float)
        dl, dh = other.dirichlet[0], other.dirichlet[1]
        low = np.ascontiguousarray(np.broadcast_to(np.atleast_1d(
            np.asarray(0.0 if dl is None else dl, dtype=float)), (blocks,)))
        high = np.ascontiguousarray(np.broadcast_to(np.atleast_1d(
            np.asarray(0.0 if dh is None else dh, dtype=float)), (blocks,)))
        deltas = getattr(other, "deltas", None)
        deltas = (np.ascontiguousarray(np.asarray(deltas, dtype=float))
                  if deltas is not None else np.zeros(0))
        h = ctypes.c_void_p()
        _lib.call("fdadi_tri_create", _lib.ptr(sup), _lib.ptr(mid), _lib.ptr(sub),
                  _lib.ptr(R), _lib.ptr(deltas), deltas.size, _lib.ptr(low), _lib.ptr(high),
                  _lib.ptr(tops), _lib.ptr(bots), int(other.axis), n, blocks,
                  int(dl is not None), int(dh is not None), _STATUS_I[top_status],
                  _STATUS_I[bottom_status], int(has_R), ctypes.byref(h))
        self._tri = h
        self.blocks = blocks
        self.axis = int(other.axis)

    def _emigrate_csr(self, other):
        data, offsets = _dia_of(other)
        N = data.shape[1]
        n1 = max(offsets) - 1
        if n1 < 2 or N % n1:
            raise ValueError("cross operator offsets do not describe a 9-point stencil")
        n0 = N // n1
        bands = np.zeros((9, N))
        for a in (1, 0, -1):
            for b in (1, 0, -1):
                k = (1 - a) * 3 + (1 - b)
                bands[k] = _row_band(data, offsets, a * n1 + b, N)
        R = getattr(other, "R", None)
        R = np.zeros(N) if R is None else np.ascontiguousarray(R, dtype=float)
        h = ctypes.c_void_p()
        _lib.call("fdadi_cross_create", _lib.ptr(np.ascontiguousarray(bands)), _lib.ptr(R),
                  n0, n1, ctypes.byref(h))
        self._csr = h
        self.axis = 1
        self.blocks = int(getattr(other, "blocks", 1))

    def immigrate(self, tag=""):
        """Host copy in the CPU BandedOperator layout (BandedOperatorGPU.pyx:166-372)."""
        if self.is_mixed_derivative:
            return self._immigrate_csr()
        return self._immigrate_tri()

    def _immigrate_tri(self):
        rows, blocks, axis, hl, hh, ts, bs, hr, nd = self._tri_info()
        bl = rows // blocks
        sup, mid, sub, R = (np.zeros(rows) for _ in range(4))
        low, high, tops, bots = (np.zeros(blocks) for _ in range(4))
        deltas = np.zeros(nd)
        _lib.call("fdadi_tri_read", self._tri, _lib.ptr(sup), _lib.ptr(mid), _lib.ptr(sub),
                  _lib.ptr(R), _lib.ptr(low), _lib.ptr(high), _lib.ptr(tops), _lib.ptr(bots),
                  _lib.ptr(deltas) if nd else None)
        offsets = [1, 0, -1]
        bands = [sup, mid, sub]
        if _STATUS_S[ts] == CAN_FOLD:
            z = np.zeros(rows)
            z[np.arange(blocks) * bl] = tops
            offsets.insert(0, 2)
            bands.insert(0, z)
        if _STATUS_S[bs] == CAN_FOLD:
            z = np.zeros(rows)
            z[np.arange(blocks) * bl + bl - 1] = bots
            offsets.append(-2)
            bands.append(z)
        data = _dia_from_rows(bands, offsets, rows)
        B = HostBandedOperator(data, offsets, R if hr else np.zeros(rows), axis=axis,
                               blocks=blocks, derivative=self.derivative, order=self.order,
                               deltas=deltas if nd else self.deltas)
        B.dirichlet = [tuple(low) if hl else None, tuple(high) if hh else None]
        B.top_fold_status = _STATUS_S[ts]
        B.bottom_fold_status = _STATUS_S[bs]
        B.top_factors = tops if B.top_fold_status == FOLDED else None
        B.bottom_factors = bots if B.bottom_fold_status == FOLDED else None
        return B

    def _immigrate_csr(self):
        n0, n1 = _lib.c_i64(), _lib.c_i64()
        _lib.call("fdadi_cross_info", self._csr, ctypes.byref(n0), ctypes.byref(n1))
        n0, n1 = n0.value, n1.value
        N = n0 * n1
        bands = np.zeros((9, N))
        R = np.zeros(N)
        _lib.call("fdadi_cross_read", self._csr, _lib.ptr(bands), _lib.ptr(R))
        offsets = [a * n1 + b for a in (1, 0, -1) for b in (1, 0, -1)]
        data = _dia_from_rows(list(bands), offsets, N)
        B = HostBandedOperator(data, offsets, None, axis=1, blocks=self.blocks,
                               derivative=self.derivative, order=self.order)
        B.is_mixed_derivative = True
        return B

    # -- folding ------------------------------------------------------------------------
    def diagonalize(self):
        """Fold this operator to make it tridiagonal."""
        if self._tri is None:
            raise RuntimeError("Diagonalize called but C++ tridiag is NULL. CSR? Mixed(%s)"
                               % self.is_mixed_derivative)
        _lib.call("fdadi_tri_diagonalize", self._tri, _lib.default_stream())

    def undiagonalize(self):
        if self._tri is None:
            raise RuntimeError("Undiagonalize called but C++ tridiag is NULL. CSR? Mixed(%s)"
                               % self.is_mixed_derivative)
        _lib.call("fdadi_tri_undiagonalize", self._tri, _lib.default_stream())

    def is_folded(self):
        return FOLDED in (self.top_fold_status, self.bottom_fold_status)

    def is_foldable(self):
        return CAN_FOLD in (self.top_fold_status, self.bottom_fold_status)

    def _no_mixed(self):
        if self.is_mixed_derivative:
            raise ValueError("Operation not supported with mixed operator.")

    def fold_top(self, unfold=False):
        self._no_mixed()
        _lib.call("fdadi_tri_fold_top", self._tri, int(bool(unfold)), _lib.default_stream())

    def fold_bottom(self, unfold=False):
        self._no_mixed()
        _lib.call("fdadi_tri_fold_bottom", self._tri, int(bool(unfold)), _lib.default_stream())

    # -- apply / solve --------------------------------------------------------------------
    def apply_(self, sa_V, overwrite):
        """Apply in place on the device array (or on a copy unless ``overwrite``)."""
        sa_U = sa_V if overwrite else sa_V.copy(True)
        n = sa_U.size
        tmp = _scratch_buffer(n)
        s = _lib.default_stream()
        if self._tri is not None:
            _lib.call("fdadi_tri_apply", self._tri, sa_U.p, tmp.ptr, n, s)
        else:
            _lib.call("fdadi_cross_apply", self._csr, sa_U.p, tmp.ptr, n, s)
        _lib.call("fdadi_array_copy", sa_U.p, tmp.ptr, n, s)
        return sa_U

    def apply(self, V, overwrite=False):
        """Wrapper for CPU based data."""
        sa_V = SizedArrayPtr(np.asarray(V), "sa_V apply")
        return self.apply_(sa_V, True).to_numpy()

    def solve_(self, sa_V, overwrite):
        sa_U = sa_V if overwrite else sa_V.copy(True)
        if self.is_mixed_derivative:
            raise NotImplementedError("Cross-derivative solve is not part of any scheme "
                                      "(the reference's csrsv path is untrustworthy).")
        _lib.call("fdadi_tri_solve", self._tri, sa_U.p, sa_U.size, _lib.default_stream())
        return sa_U

    def solve(self, V, overwrite=False):
        sa_V = SizedArrayPtr(np.asarray(V), "sa_V solve")
        return self.solve_(sa_V, True).to_numpy()

    def enable_residual(self, state):
        """If state is false, the residual is not used when applying or solving."""
        if self._tri is not None:
            _lib.call("fdadi_tri_set_flags", self._tri, -1, -1, int(bool(state)))

    # -- algebra ---------------------------------------------------------------------------
    def mul_scalar_from_host(self, v, inplace=False):
        val = SizedArrayPtr(np.atleast_1d(float(v)))
        if inplace:
            self.vectorized_scale(val)
            return self
        return self * val

    def __mul__(self, val):
        B = self.copy()
        B.__imul__(val)
        return B

    def __imul__(self, val):
        if not isinstance(val, SizedArrayPtr):
            val = SizedArrayPtr(np.atleast_1d(float(val)))
        self.vectorized_scale(val)
        return self

    def add(self, val, inplace=False):
        if inplace:
            self += val
            return self
        return self + val

    def __add__(self, val):
        B = self.copy()
        if isinstance(val, BandedOperator):
            B.add_operator(val)
        else:
            B.add_scalar_from_host(val)
        return B

    def __iadd__(self, val):
        if isinstance(val, BandedOperator):
            self.add_operator(val)
        else:
            self.add_scalar_from_host(val)
        return self

    def add_operator(self, other):
        """Add a second BandedOperator to this one (residuals add too)."""
        if self.is_mixed_derivative:
            raise NotImplementedError("No add to mixed operator")
        elif other.is_mixed_derivative:
            raise NotImplementedError("No add mixed operator to this one")
        elif self.axis != other.axis:
            raise ValueError("Both operators must operate on the same axis."
                             " (%s != %s)" % (self.axis, other.axis))
        elif self.operator_rows != other.operator_rows:
            raise ValueError("Both operators must have the same length")
        elif self.is_folded():
            raise NotImplementedError("No add to diagonalized operator.")
        elif other.is_folded():
            raise NotImplementedError("No add diagonalized operator.")
        _lib.call("fdadi_tri_add_operator", self._tri, other._tri, _lib.default_stream())

    def add_scalar_from_host(self, other):
        """Add a scalar to the main diagonal (Dirichlet rows skipped)."""
        if self._tri is None:
            raise RuntimeError("add_scalar called but C++ tridiag is NULL. CSR? Mixed(%s)"
                               % self.is_mixed_derivative)
        _lib.call("fdadi_tri_add_scalar", self._tri, float(other), _lib.default_stream())

    def add_scalar(self, other, index=0):
        """Add the device scalar other[index] to the main diagonal."""
        if self._tri is None:
            raise RuntimeError("add_scalar called but C++ tridiag is NULL. CSR? Mixed(%s)"
                               % self.is_mixed_derivative)
        if not 0 <= index < other.size:
            raise IndexError("scalar index %d out of range" % index)
        p = ctypes.c_void_p(other.p.value + 8 * int(index))
        _lib.call("fdadi_tri_add_scalar_dev", self._tri, p, _lib.default_stream())

    def vectorized_scale(self, vector):
        """Row-scale the operator (and R) by the device vector, tiled."""
        if self.is_mixed_derivative:
            _lib.call("fdadi_cross_scale", self._csr, vector.p, vector.size,
                      _lib.default_stream())
        else:
            _lib.call("fdadi_tri_scale", self._tri, vector.p, vector.size,
                      _lib.default_stream())

    def vectorized_scale_from_host(self, vector):
        v = SizedArrayPtr(np.asarray(vector, dtype=float), "Vectorized scale v")
        self.vectorized_scale(v)
        _lib.sync()
        del v


def for_vector(v, blocks, derivative, axis, barrier):
    """Heston 1-D operator for mesh ``v`` repeated over ``blocks`` lines
    (BandedOperatorGPU.pyx:599-618, _TriBandedOperatorGPU.cu:1107-1182)."""
    v = np.ascontiguousarray(v, dtype=float)
    B = BandedOperator(tag='for_vector')
    h = ctypes.c_void_p()
    _lib.call("fdadi_tri_for_vector", _lib.ptr(v), v.size, int(blocks), int(derivative),
              int(axis), int(bool(barrier)), _lib.default_stream(), ctypes.byref(h))
    B._tri = h
    B.blocks = int(blocks)
    B.is_mixed_derivative = False
    B.axis = int(axis)
    B.order = 2
    B.derivative = int(derivative)
    B.deltas = None
    return B


def mixed_for_vector(v0, v1):
    """d2 / dv0 dv1 on the (len(v0), len(v1)) grid (BandedOperatorGPU.pyx:621-641)."""
    v0 = np.ascontiguousarray(v0, dtype=float)
    v1 = np.ascontiguousarray(v1, dtype=float)
    B = BandedOperator(tag='mixed_for_vector')
    h = ctypes.c_void_p()
    _lib.call("fdadi_cross_for_vector", _lib.ptr(v0), v0.size, _lib.ptr(v1), v1.size,
              _lib.default_stream(), ctypes.byref(h))
    B._csr = h
    B.is_mixed_derivative = True
    B.derivative = 1
    B.axis = 1
    B.order = 2
    B.deltas = None
    B.blocks = v0.size
    return B


def cublas_to_scipy(B):
    """Row-indexed bands -> scipy DIA layout, in place on B.data
    (BandedOperatorGPU.pyx:651-662)."""
    for row, o in enumerate(B.offsets):
        if o > 0:
            B.data[row, o:] = B.data[row, :-o].copy()
            B.data[row, :o] = 0
        if o < 0:
            B.data[row, :o] = B.data[row, -o:].copy()
            B.data[row, o:] = 0


def scipy_to_cublas(B):
    """scipy DIA layout -> row-indexed bands (BandedOperatorGPU.pyx:665-676)."""
    for row, o in enumerate(B.offsets):
        if o > 0:
            B.data[row, :-o] = B.data[row, o:].copy()
            B.data[row, -o:] = 0
        if o < 0:
            B.data[row, -o:] = B.data[row, :o].copy()
            B.data[row, :-o] = 0
