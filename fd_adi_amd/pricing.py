"""Semi-analytic Heston European call prices on the device.

The reference checks its finite-difference prices against a semi-analytic
COS pricer (FiniteDifference/heston.py:490-693).  ``heston_call`` prices with
the library's device kernel instead (Lewis's single-integral formula over the
"little trap" characteristic function, FP64 Gauss-Legendre quadrature; one
64-lane wave per option).  Rows are S0, K, r, v0, T, kappa, theta, sigma, rho.
"""

import ctypes

import numpy as np

from . import _lib
from .SizedArrayPtr import SizedArrayPtr

ROW = ("S0", "K", "r", "v0", "T", "kappa", "theta", "sigma", "rho")


def heston_call_device(rows, n, out, stream=None):
    """Prices of the ``n`` rows of device array ``rows`` [n x 9] into device
    array ``out`` [n] (asynchronous on ``stream``)."""
    _lib.call("fdadi_heston_cf_call", rows.p, int(n), out.p,
              _lib.default_stream() if stream is None else stream)


def heston_call(S0, K, r, v0, T, kappa, theta, sigma, rho):
    """Call prices for arrays (or scalars) of parameters, broadcast together;
    returns a float64 array of the broadcast shape (a float for scalars)."""
    cols = np.broadcast_arrays(*[np.asarray(a, dtype=float) for a in
                                 (S0, K, r, v0, T, kappa, theta, sigma, rho)])
    shape = cols[0].shape
    rows = np.ascontiguousarray(np.stack([c.ravel() for c in cols], axis=1))
    n = rows.shape[0]
    if n == 0:
        return np.zeros(shape)
    d_rows = SizedArrayPtr(rows)
    d_out = SizedArrayPtr().alloc(n)
    heston_call_device(d_rows, n, d_out)
    out = d_out.to_numpy().reshape(shape)
    return float(out) if out.ndim == 0 else out


def gauss_legendre(n):
    """The host Gauss-Legendre rule of the device pricer: (nodes, weights)."""
    x, w = np.zeros(n), np.zeros(n)
    _lib.call("fdadi_gauss_legendre", int(n), x.ctypes.data_as(ctypes.c_void_p),
              w.ctypes.data_as(ctypes.c_void_p))
    return x, w
