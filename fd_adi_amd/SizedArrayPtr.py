"""Device array handle -- drop-in for FiniteDifference/SizedArrayPtr.pyx.

A ``SizedArrayPtr`` owns (or views) an FP64 device buffer allocated by
libfdadi.  The methods keep the reference's names and meaning
(SizedArrayPtr.pyx:23-182): H2D on construction from a numpy array,
``to_numpy`` (D2H), deep/shallow ``copy``, ``copy_from`` and the in-place
elementwise operators.  A shallow copy is a view that becomes invalid when its
owner is released (SizedArrayPtr.pyx:160-178, tests/test_SizedArray.py:33-38):
using it then raises RuntimeError.

Work is queued on the library stream; ``to_numpy`` synchronises.
"""

import ctypes

import numpy as np

from . import _lib

_OP = {"pluseq": 0, "minuseq": 1, "minuseq_over2": 2, "timeseq": 3}


class _Buffer(object):
    """A device allocation with an explicit lifetime flag."""

    __slots__ = ("ptr", "size", "alive")

    def __init__(self, size):
        p = ctypes.c_void_p()
        _lib.call("fdadi_array_alloc", int(size), ctypes.byref(p))
        self.ptr = p
        self.size = int(size)
        self.alive = True

    def release(self):
        if self.alive:
            self.alive = False
            if self.ptr:
                _lib.lib.fdadi_array_free(self.ptr)
            self.ptr = ctypes.c_void_p()


class SizedArrayPtr(object):

    def __init__(self, a=None, tag="Unknown"):
        self.tag = tag
        self._buf = None
        self._owner = True
        self.shape = (0,)
        if a is not None:
            self.from_numpy(np.asarray(a), tag)

    # -- construction ---------------------------------------------------------------
    def alloc(self, sz, tag="Unknown"):
        """Allocate ``sz`` uninitialised doubles (SizedArrayPtr.pyx:36-46)."""
        self._drop()
        self._buf = _Buffer(sz)
        self._owner = True
        self.shape = (int(sz),)
        self.tag = tag
        return self

    def from_numpy(self, a, tag="Unknown"):
        """Upload ``a`` (any shape, FP64) (to_SizedArray, SizedArrayPtr.pyx:320-326)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.alloc(a.size, tag)
        self.shape = a.shape if a.ndim else (1,)
        _lib.call("fdadi_array_h2d", self._buf.ptr, _lib.ptr(a), a.size, _lib.default_stream())
        return self

    @classmethod
    def from_device(cls, ptr, shape, owner_ref=None, tag="view"):
        """Wrap a device pointer owned elsewhere (e.g. an engine's arena)."""
        S = cls(tag=tag)
        buf = _Buffer.__new__(_Buffer)
        buf.ptr = ctypes.c_void_p(ptr if isinstance(ptr, int) else ptr.value)
        buf.size = int(np.prod(shape))
        buf.alive = True
        S._buf = buf
        S._owner = False
        S._keep = owner_ref
        S.shape = tuple(shape)
        return S

    # -- properties -----------------------------------------------------------------
    @property
    def size(self):
        return 0 if self._buf is None else self._buf.size

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def p(self):
        """Raw device pointer (ctypes.c_void_p)."""
        return self._live().ptr

    def _live(self):
        if self._buf is None or not self._buf.alive:
            raise RuntimeError("SizedArrayPtr %r: device buffer is not alive "
                               "(owner released)" % self.tag)
        return self._buf

    # -- transfers ------------------------------------------------------------------
    def to_numpy(self):
        """D2H copy (from_SizedArray, SizedArrayPtr.pyx:349-357)."""
        b = self._live()
        out = np.empty(b.size, dtype=np.float64)
        _lib.call("fdadi_array_d2h", _lib.ptr(out), b.ptr, b.size, _lib.default_stream())
        return out.reshape(self.shape)

    def copy(self, deep=True):
        """Deep copy, or a shallow view that dies with this owner."""
        b = self._live()
        if not deep:
            S = SizedArrayPtr(tag=self.tag + " (view)")
            S._buf = b
            S._owner = False
            S.shape = self.shape
            return S
        S = SizedArrayPtr().alloc(b.size, self.tag)
        S.shape = self.shape
        _lib.call("fdadi_array_copy", S._buf.ptr, b.ptr, b.size, _lib.default_stream())
        return S

    def copy_from(self, other):
        """Device-to-device copy of ``other`` into this buffer (VecArray.h:400-410)."""
        a, b = self._live(), other._live()
        if a.size != b.size:
            raise ValueError("copy_from: size mismatch (%d != %d)" % (a.size, b.size))
        _lib.call("fdadi_array_copy", a.ptr, b.ptr, a.size, _lib.default_stream())
        self.shape = other.shape
        return self

    def transpose(self, strategy=1):
        """(h, w) -> (w, h) in place semantics (VecArray.h:438-464)."""
        if len(self.shape) != 2:
            raise ValueError("transpose needs a 2-D array")
        h, w = self.shape
        b = self._live()
        tmp = _Buffer(b.size)
        _lib.call("fdadi_array_transpose", tmp.ptr, b.ptr, h, w, _lib.default_stream())
        _lib.call("fdadi_array_copy", b.ptr, tmp.ptr, b.size, _lib.default_stream())
        _lib.sync()
        tmp.release()
        self.shape = (w, h)
        return self

    # -- elementwise ----------------------------------------------------------------
    def _binop(self, other, op):
        a, b = self._live(), other._live()
        if a.size != b.size:
            raise ValueError("size mismatch (%d != %d)" % (a.size, b.size))
        _lib.call("fdadi_array_binop", a.ptr, b.ptr, a.size, _OP[op], _lib.default_stream())
        return self

    def pluseq(self, other):
        return self._binop(other, "pluseq")

    def minuseq(self, other):
        return self._binop(other, "minuseq")

    def minuseq_over2(self, other):
        """self = (self - other) / 2 (VecArray.cu minuseq_over2)."""
        return self._binop(other, "minuseq_over2")

    def timeseq(self, other):
        return self._binop(other, "timeseq")

    def _scalar_dev(self, other, i, op):
        a, s = self._live(), other._live()
        if not 0 <= i < s.size:
            raise IndexError("scalar index %d out of range" % i)
        sp = ctypes.c_void_p(s.ptr.value + 8 * int(i))
        _lib.call("fdadi_array_scalar_op_dev", a.ptr, sp, a.size, op, _lib.default_stream())
        return self

    def pluseq_scalar(self, other, i=0):
        return self._scalar_dev(other, i, 0)

    def minuseq_scalar(self, other, i=0):
        return self._scalar_dev(other, i, 1)

    def timeseq_scalar(self, other, i=0):
        return self._scalar_dev(other, i, 3)

    def _scalar_host(self, v, op):
        a = self._live()
        _lib.call("fdadi_array_scalar_op", a.ptr, float(v), a.size, op, _lib.default_stream())
        return self

    def pluseq_scalar_from_host(self, other):
        return self._scalar_host(other, 0)

    def minuseq_scalar_from_host(self, other):
        return self._scalar_host(other, 1)

    def timeseq_scalar_from_host(self, other):
        return self._scalar_host(other, 3)

    def __neg__(self):
        S = self.copy(True)
        S.timeseq_scalar_from_host(-1.0)
        return S

    # -- lifetime ---------------------------------------------------------------------
    def _drop(self):
        if self._buf is not None and self._owner:
            self._buf.release()
        self._buf = None

    def __del__(self):
        try:
            self._drop()
        except Exception:
            pass

    def __str__(self):
        return "SizedArrayPtr(%s, shape=%s)" % (self.tag, self.shape)

    __repr__ = __str__
