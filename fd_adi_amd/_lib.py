"""ctypes binding of libfdadi.so (the C ABI declared in include/fdadi.h).

The library must have been built (``python fd_adi_amd/build.py``); there is
no fallback: importing the package without it raises ImportError.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# FDADI_LIB names another build of the same library (e.g. one compiled for a
# different architecture); by default the in-tree build is loaded.
LIB_PATH = os.environ.get("FDADI_LIB") or os.path.join(_HERE, "libfdadi.so")

OK, ERR_ARG, ERR_STATE, ERR_HIP, ERR_UNSUPPORTED = 0, 1, 2, 3, 4
CANNOT_FOLD_I, CAN_FOLD_I, FOLDED_I = 0, 1, 2
SCHEME_IMPLICIT, SCHEME_DOUGLAS, SCHEME_CS, SCHEME_MCS, SCHEME_HV = 0, 1, 2, 3, 4

if not os.path.exists(LIB_PATH):
    raise ImportError("fd_adi_amd: %s is missing; build it with "
                      "`python fd_adi_amd/build.py` (no CPU fallback exists)" % LIB_PATH)

lib = ctypes.CDLL(LIB_PATH)

c_dp = ctypes.POINTER(ctypes.c_double)
c_i64 = ctypes.c_int64
c_i32 = ctypes.c_int32
c_int = ctypes.c_int
c_vp = ctypes.c_void_p
c_double = ctypes.c_double


class HestonOptionStruct(ctypes.Structure):
    """fdadi_heston_option."""
    _fields_ = [("r", c_double), ("kappa", c_double), ("theta", c_double),
                ("sigma", c_double), ("rho", c_double), ("dt", c_double),
                ("nS", c_i64), ("nv", c_i64), ("barrier", c_i32), ("pad", c_i32)]


_P = ctypes.POINTER
_SIGS = {
    "fdadi_last_error": (ctypes.c_char_p, []),
    "fdadi_version": (c_int, [_P(c_int), _P(c_int)]),
    "fdadi_device_count": (c_int, [_P(c_int)]),
    "fdadi_set_device": (c_int, [c_int]),
    "fdadi_stream_create": (c_int, [_P(c_vp)]),
    "fdadi_stream_destroy": (c_int, [c_vp]),
    "fdadi_stream_sync": (c_int, [c_vp]),
    "fdadi_device_sync": (c_int, []),
    "fdadi_array_alloc": (c_int, [c_i64, _P(c_vp)]),
    "fdadi_array_free": (c_int, [c_vp]),
    "fdadi_array_h2d": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "fdadi_array_d2h": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "fdadi_array_copy": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "fdadi_array_fill": (c_int, [c_vp, c_double, c_i64, c_vp]),
    "fdadi_array_binop": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp]),
    "fdadi_array_scalar_op": (c_int, [c_vp, c_double, c_i64, c_int, c_vp]),
    "fdadi_array_scalar_op_dev": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp]),
    "fdadi_array_transpose": (c_int, [c_vp, c_vp, c_i64, c_i64, c_vp]),
    "fdadi_arena_create": (c_int, [c_i64, _P(c_vp)]),
    "fdadi_arena_alloc": (c_int, [c_vp, c_i64, _P(c_vp)]),
    "fdadi_arena_used": (c_int, [c_vp, _P(c_i64)]),
    "fdadi_arena_reset": (c_int, [c_vp]),
    "fdadi_arena_destroy": (c_int, [c_vp]),
    "fdadi_tri_create": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp,
                                 c_int, c_i64, c_i64, c_int, c_int, c_int, c_int, c_int,
                                 _P(c_vp)]),
    "fdadi_tri_for_vector": (c_int, [c_vp, c_i64, c_i64, c_int, c_int, c_int, c_vp, _P(c_vp)]),
    "fdadi_tri_copy": (c_int, [c_vp, c_vp, _P(c_vp)]),
    "fdadi_tri_destroy": (c_int, [c_vp]),
    "fdadi_tri_info": (c_int, [c_vp, _P(c_i64), _P(c_i64), _P(c_int), _P(c_int), _P(c_int),
                               _P(c_int), _P(c_int), _P(c_int), _P(c_i64)]),
    "fdadi_tri_set_flags": (c_int, [c_vp, c_int, c_int, c_int]),
    "fdadi_tri_read": (c_int, [c_vp] + [c_vp] * 9),
    "fdadi_tri_apply": (c_int, [c_vp, c_vp, c_vp, c_i64, c_vp]),
    "fdadi_tri_solve": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "fdadi_tri_add_operator": (c_int, [c_vp, c_vp, c_vp]),
    "fdadi_tri_add_scalar": (c_int, [c_vp, c_double, c_vp]),
    "fdadi_tri_add_scalar_dev": (c_int, [c_vp, c_vp, c_vp]),
    "fdadi_tri_scale": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "fdadi_tri_scale_scalar": (c_int, [c_vp, c_double, c_vp]),
    "fdadi_tri_fold_top": (c_int, [c_vp, c_int, c_vp]),
    "fdadi_tri_fold_bottom": (c_int, [c_vp, c_int, c_vp]),
    "fdadi_tri_diagonalize": (c_int, [c_vp, c_vp]),
    "fdadi_tri_undiagonalize": (c_int, [c_vp, c_vp]),
    "fdadi_cross_for_vector": (c_int, [c_vp, c_i64, c_vp, c_i64, c_vp, _P(c_vp)]),
    "fdadi_cross_create": (c_int, [c_vp, c_vp, c_i64, c_i64, _P(c_vp)]),
    "fdadi_cross_copy": (c_int, [c_vp, c_vp, _P(c_vp)]),
    "fdadi_cross_destroy": (c_int, [c_vp]),
    "fdadi_cross_info": (c_int, [c_vp, _P(c_i64), _P(c_i64)]),
    "fdadi_cross_read": (c_int, [c_vp, c_vp, c_vp]),
    "fdadi_cross_apply": (c_int, [c_vp, c_vp, c_vp, c_i64, c_vp]),
    "fdadi_cross_scale": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "fdadi_cross_scale_scalar": (c_int, [c_vp, c_double, c_vp]),
    "fdadi_heston_coefficients": (c_int, [c_int, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp]),
    "fdadi_hb_create": (c_int, [c_i32, c_vp, c_vp, c_vp, c_vp, _P(c_vp)]),
    "fdadi_hb_destroy": (c_int, [c_vp]),
    "fdadi_hb_configure": (c_int, [c_vp, c_int, c_double, c_double, c_vp]),
    "fdadi_hb_configure_dev": (c_int, [c_vp, c_int, c_vp, c_vp, c_vp]),
    "fdadi_hb_update": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp]),
    "fdadi_hb_update_dev": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "fdadi_hb_call_payoff": (c_int, [c_vp, c_vp, c_vp]),
    "fdadi_hb_read_bicubic": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "fdadi_hb_state": (c_int, [c_vp, _P(c_vp), _P(c_i64)]),
    "fdadi_hb_offsets": (c_int, [c_vp, c_vp]),
    "fdadi_hb_step": (c_int, [c_vp, c_i32, c_vp]),
    "fdadi_hb_step_timed": (c_int, [c_vp, c_i32, c_vp, c_vp, c_i32, _P(c_i32)]),
    "fdadi_hb_capture": (c_int, [c_vp, c_i32, c_vp]),
    "fdadi_hb_replay": (c_int, [c_vp, c_vp]),
    "fdadi_hb_capture_solve": (c_int, [c_vp, c_int, c_double, c_i32, c_i32, c_vp]),
    "fdadi_hb_set_shape": (c_int, [c_vp, c_int, c_int, c_int]),
    "fdadi_hb_get_shape": (c_int, [c_vp, c_int, _P(c_int), _P(c_int)]),
    "fdadi_hb_launches_per_step": (c_int, [c_vp, _P(c_int)]),
    "fdadi_mlp_nparam": (c_int, [_P(c_i64)]),
    "fdadi_mlp_create": (c_int, [c_i32, c_vp, c_i64, c_vp, c_vp, c_double, c_double, _P(c_vp)]),
    "fdadi_mlp_create_features": (c_int, [c_i32, c_i32, c_vp, c_i64, c_vp, c_vp, c_double,
                                          c_double, _P(c_vp)]),
    "fdadi_mlp_destroy": (c_int, [c_vp]),
    "fdadi_mlp_get_params": (c_int, [c_vp, c_vp, c_vp]),
    "fdadi_mlp_set_params": (c_int, [c_vp, c_vp, c_vp]),
    "fdadi_mlp_grad": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp]),
    "fdadi_mlp_train_step": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i32, ctypes.c_float, c_vp]),
    "fdadi_mlp_loss": (c_int, [c_vp, c_vp]),
    "fdadi_mlp_predict": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp]),
    "fdadi_mlp_predict_noalloc": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp]),
    "fdadi_mlp_gemm_nt": (c_int, [c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp]),
    "fdadi_heston_cf_call": (c_int, [c_vp, c_i64, c_vp, c_vp]),
    "fdadi_forest_create": (c_int, [c_i32, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                    c_double, _P(c_vp)]),
    "fdadi_forest_destroy": (c_int, [c_vp]),
    "fdadi_forest_predict": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp]),
    "fdadi_serve_combine": (c_int, [c_vp, c_i64, c_vp, c_i32, c_vp, c_vp, c_vp, c_double, c_vp,
                                    c_vp, c_vp, c_vp]),
    "fdadi_serve_cf_rows": (c_int, [c_vp, c_i64, c_vp, c_vp]),
    "fdadi_serve_gate": (c_int, [c_vp, c_vp, c_i64, c_double, c_vp, c_vp, c_vp, _P(c_i32),
                                 c_vp]),
    "fdadi_serve_label_inputs": (c_int, [c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_i32, c_vp,
                                         c_vp, c_vp, c_vp, c_vp, c_vp]),
    "fdadi_serve_scatter": (c_int, [c_vp, c_vp, c_i32, c_vp, c_vp]),
    "fdadi_serve_max_rows": (c_int, [c_vp, c_i32, c_i64, c_vp, c_vp]),
    "fdadi_gauss_legendre": (c_int, [c_i32, c_vp, c_vp]),
}

def _declare(table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


_declare(_SIGS)

EXPORTED = sorted(_SIGS)


def declare_extra(table):
    """Declare additional entry points (used by the surrogate module)."""
    _declare(table)
    EXPORTED.extend(sorted(table))
    EXPORTED.sort()


def last_error():
    msg = lib.fdadi_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(status, what=""):
    """Raise the Python exception matching an fdadi status code."""
    if status == OK:
        return
    msg = last_error()
    if what:
        msg = "%s: %s" % (what, msg)
    if status == ERR_ARG:
        raise ValueError(msg)
    if status == ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def call(name, *args):
    check(getattr(lib, name)(*args), name)


def ptr(a):
    """Host pointer of a contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(c_vp)


_default_stream = None


def default_stream():
    """The library stream used by the drop-in classes (created on first use)."""
    global _default_stream
    if _default_stream is None:
        s = c_vp()
        call("fdadi_stream_create", ctypes.byref(s))
        _default_stream = s
    return _default_stream


def sync():
    call("fdadi_stream_sync", default_stream())
