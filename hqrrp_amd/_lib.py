"""ctypes binding of libhqrrp_amd.so (the C ABI declared in include/hqrrp_amd.h).

The library is built in-tree (``make -C hqrrp_amd/csrc``); importing a
symbol when it is missing raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhqrrp_amd.so")

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_vp = C.c_void_p


class Trace(C.Structure):
    """Mirror of ``hqrrp_trace`` (device pointers or NULL)."""

    _fields_ = [
        ("sketch_cols", _vp),
        ("y0", _vp),
        ("y", _vp),
        ("g", _vp),
    ]


class Options(C.Structure):
    """Mirror of ``hqrrp_options``."""

    _fields_ = [
        ("nb", C.c_int),
        ("pp", C.c_int),
        ("panel_pivoting", C.c_int),
        ("sketch_prec", C.c_int),
        ("omega_kind", C.c_int),
        ("seed", C.c_uint),
        ("omega", _vp),
        ("max_rank", C.c_int),
        ("lookahead", C.c_int),
        ("trace", C.POINTER(Trace)),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "dgeqp4": (None, [_i, _i, _vp, _i, _vp, _vp, _vp, _i, _i]),
    "NoFLA_HQRRP_WY_blk_var4": (C.c_int, [C.c_int, C.c_int, _vp, C.c_int, _vp,
                                          _vp, C.c_int, C.c_int, C.c_int]),
    "hqrrp_ut_blk_var2": (C.c_int, [C.c_int, C.c_int, _vp, C.c_int, _vp, _vp,
                                    C.c_int, C.c_int, C.c_int]),
    "hqrrp_ut_factor_dev": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int,
                                      _vp, _vp, C.POINTER(Options)]),
    "hqrrp_default_options": (None, [C.POINTER(Options)]),
    "hqrrp_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "hqrrp_destroy": (C.c_int, [_vp]),
    "hqrrp_set_stream": (C.c_int, [_vp, _vp]),
    "hqrrp_reset_stream": (C.c_int, [_vp]),
    "hqrrp_set_panel_alg": (C.c_int, [_vp, C.c_int]),
    "hqrrp_synchronize": (C.c_int, [_vp]),
    "hqrrp_phase_count": (C.c_int, []),
    "hqrrp_phase_name": (C.c_char_p, [C.c_int]),
    "hqrrp_profile_enable": (C.c_int, [_vp, C.c_int]),
    "hqrrp_profile_phases": (C.c_int, [_vp, C.c_uint]),
    "hqrrp_profile_read": (C.c_int, [_vp, C.c_int, _d, C.POINTER(C.c_longlong),
                                     _d, _d]),
    "hqrrp_factor_dev": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, _vp,
                                   _vp, C.POINTER(Options)]),
    "hqrrp_reference_omega": (None, [C.c_int, C.c_int, _vp, C.c_uint]),
    "hqrrp_reference_omega_carry": (None, [C.c_int, C.c_int, _vp, C.c_uint,
                                           C.POINTER(C.c_int),
                                           C.POINTER(C.c_double)]),
    "hqrrp_omega_carry_get": (None, [C.POINTER(C.c_int),
                                     C.POINTER(C.c_double)]),
    "hqrrp_omega_carry_set": (None, [C.c_int, C.c_double]),
    "hqrrp_sketch_qrcp_dev": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int,
                                        C.c_int, _vp, _vp]),
    "hqrrp_panel_qrcp_dev": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int,
                                       C.c_int, _vp, _vp, _vp, C.c_int]),
    "hqrrp_apply_qt_left_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp,
                                          C.c_int, _vp, C.c_int, _vp, C.c_int]),
    "hqrrp_downdate_y_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                       _vp, C.c_int, _vp, C.c_int, _vp,
                                       C.c_int, _vp, C.c_int, _vp, C.c_int]),
    "hqrrp_dgemm_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_double, _vp, C.c_int, _vp,
                                  C.c_int, C.c_double, _vp, C.c_int]),
    "hqrrp_geqrf_dev": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, _vp]),
    "hqrrp_geqp3_dev": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, _vp,
                                 _vp]),
    "hqrrp_geqp3_noalloc_dev": (C.c_int, [_vp, C.c_int, C.c_int, _vp,
                                         C.c_int, _vp, _vp]),
    "hqrrp_orgqr_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp,
                                  C.c_int, _vp, _vp, C.c_int]),
    "hqrrp_ormqr_dev": (C.c_int, [_vp, C.c_char, C.c_char, C.c_int, C.c_int,
                                  C.c_int, _vp, C.c_int, _vp, _vp, C.c_int]),
    "hqrrp_qr_fixed_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp,
                                     C.c_int, _vp, _vp, C.c_int, _vp,
                                     C.c_int]),
    "hqrrp_batch_create": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, C.c_int,
                                     C.c_int, C.POINTER(_vp)]),
    "hqrrp_batch_get_nb": (C.c_int, [_vp, _vp]),
    "hqrrp_batch_factor_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "hqrrp_batch_destroy": (C.c_int, [_vp]),
    "hqrrp_batch_profile": (C.c_int, [_vp, C.c_int]),
    "hqrrp_batch_tune": (C.c_int, [_vp, C.c_int, C.c_int]),
    "hqrrp_batch_tune_table": (C.c_int, [_vp, _vp]),
    "hqrrp_batch_profile_read": (C.c_int, [_vp, C.c_int, _d]),
    "hqrrp_batch_phase_count": (C.c_int, []),
    "hqrrp_batch_phase_name": (C.c_char_p, [C.c_int]),
    "hqrrp_qr_fixed_adjoint_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int,
                                             _vp, C.c_int, _vp, _vp, C.c_int,
                                             _vp, C.c_int, _vp, C.c_int, _vp,
                                             C.c_int, _vp, C.c_int]),
}

_LIB = None


def lib():
    """Load the library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C hqrrp_amd/csrc` "
                "or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class HQRRPError(RuntimeError):
    pass


STATUS = {0: "OK", -1: "illegal argument", -2: "allocation failed",
          -3: "HIP runtime error", -4: "kernel failure (barrier timeout)",
          -5: "unsupported parameters"}


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise HQRRPError(f"{what} failed: {rc} ({STATUS.get(rc, '?')})")
