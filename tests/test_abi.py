"""The C-ABI library loads and exports every symbol include/fdadi.h declares
(CPU only: no compute call is made)."""

import ctypes
import os
import re

from fd_adi_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(fdadi_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_symbols_exported():
    names = declared()
    assert len(names) > 50
    missing = [n for n in sorted(names) if not hasattr(_lib.lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert declared() <= set(_lib.EXPORTED)


def test_error_channel_without_device():
    # An argument error is reported through the status code and fdadi_last_error.
    st = _lib.lib.fdadi_tri_for_vector(None, 3, 1, 1, 0, 0, None, ctypes.byref(ctypes.c_void_p()))
    assert st == _lib.ERR_ARG
    assert "null" in _lib.last_error() or "short" in _lib.last_error()
    st = _lib.lib.fdadi_array_binop(None, None, 10, 7, None)
    assert st == _lib.ERR_ARG


def test_gauss_legendre_rule_matches_numpy():
    """The host quadrature rule of the device Heston pricer (no GPU call)."""
    import numpy as np
    from fd_adi_amd.pricing import gauss_legendre
    for n in (1, 2, 7, 64, 512):
        x, w = gauss_legendre(n)
        X, W = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(x, X, rtol=0, atol=1e-15)
        np.testing.assert_allclose(w, W, rtol=0, atol=1e-14)
