"""The C ABI: libhqrrp_amd.so loads, exports every function that
include/hqrrp_amd.h declares, and its host-side logic (argument checks,
workspace query, Omega generation) behaves like the reference's -- all
without touching a GPU."""
import ctypes as C
import os
import re

import numpy as np

import hqrrp_amd
from hqrrp_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include",
                      "hqrrp_amd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(",
                       src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_header_symbols_exported():
    L = C.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert "dgeqp4" in names and "NoFLA_HQRRP_WY_blk_var4" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # The Python binding covers every declared function too.
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_default_options_are_reference_defaults():
    o = _lib.Options()
    hqrrp_amd.lib().hqrrp_default_options(C.byref(o))
    assert (o.nb, o.pp, o.panel_pivoting, o.seed) == (64, 10, 1, 12)
    assert (o.sketch_prec, o.omega_kind, o.max_rank) == (0, 0, 0)
    assert o.lookahead == 1


def _call_dgeqp4(m, n, lda, lwork):
    A = np.zeros(max(1, lda * max(n, 1)))
    jp = np.zeros(max(1, n), np.int32)
    tau = np.zeros(max(1, n))
    work = np.zeros(max(1, lwork, 4))
    info = C.c_int(99)
    hqrrp_amd.lib().dgeqp4(C.byref(C.c_int(m)), C.byref(C.c_int(n)),
                           C.c_void_p(A.ctypes.data), C.byref(C.c_int(lda)),
                           C.c_void_p(jp.ctypes.data), C.c_void_p(tau.ctypes.data),
                           C.c_void_p(work.ctypes.data), C.byref(C.c_int(lwork)),
                           C.byref(info))
    return info.value, work[0]


def test_dgeqp4_argument_errors():
    # Same info codes as the reference (NoFLA_HQRRP_WY_blk_var4.c:149-182).
    assert _call_dgeqp4(-1, 5, 1, 100)[0] == -1
    assert _call_dgeqp4(5, -1, 5, 100)[0] == -2
    assert _call_dgeqp4(5, 5, 4, 100)[0] == -4
    assert _call_dgeqp4(5, 5, 5, 15)[0] == -8  # lwork < 3n+1


def test_dgeqp4_workspace_query_and_quick_return():
    info, w0 = _call_dgeqp4(10, 8, 10, -1)
    assert info == 0 and w0 == 2 * 8 + 9 * 32
    info, w0 = _call_dgeqp4(0, 8, 1, 100)
    assert info == 0 and w0 == 1.0


def test_driver_argument_errors():
    L = hqrrp_amd.lib()
    buf = np.zeros(4)
    ip = np.zeros(4, np.int32)
    vp = C.c_void_p(buf.ctypes.data)
    assert L.NoFLA_HQRRP_WY_blk_var4(-1, 2, vp, 1, C.c_void_p(ip.ctypes.data),
                                     vp, 64, 10, 1) == -1
    assert L.NoFLA_HQRRP_WY_blk_var4(4, 2, vp, 3, C.c_void_p(ip.ctypes.data),
                                     vp, 64, 10, 1) == -1
    # Quick return for an empty matrix, as the reference.
    assert L.NoFLA_HQRRP_WY_blk_var4(0, 2, vp, 1, C.c_void_p(ip.ctypes.data),
                                     vp, 64, 10, 1) == 0


def _gpu_present():
    try:
        return C.CDLL("libamdhip64.so").hipGetDeviceCount is not None and \
            _device_count() > 0
    except OSError:
        return False


def _device_count():
    n = C.c_int(0)
    rc = C.CDLL("libamdhip64.so").hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def test_dgeqp4_device_failure_is_not_an_argument_code():
    """Without a usable device the factorisation cannot run: dgeqp4 reports
    HQRRP_INFO_DEVICE - status (> 0), never one of LAPACK's negative
    argument codes (info = -5 would read as 'jpvt illegal')."""
    import pytest

    if _gpu_present():
        pytest.skip("a device is present")
    info, _ = _call_dgeqp4(1200, 1000, 1200, 128 * 1000)
    assert info == 1000 - (-3), info  # HQRRP_ERR_HIP


def test_omega_carry_host_functions():
    """hqrrp_reference_omega_carry against the oracle's restatement of the
    reference's generator statics (itself pinned against the compiled
    reference in test_oracle.py), and the process-wide state's get/set."""
    import oracle

    r, m = 21, 301  # r * m odd: the draw leaves a cached variate
    G1, c1 = hqrrp_amd.reference_omega(r, m, carry=(0, 0.0))
    assert np.array_equal(G1, hqrrp_amd.reference_omega(r, m))
    assert np.array_equal(G1, oracle.omega(r, m))
    H1, d1 = oracle.omega(r, m, carry=(0, 0.0))
    assert c1 == d1 and c1[0] == 1
    G2, c2 = hqrrp_amd.reference_omega(r, m, carry=c1)
    H2, d2 = oracle.omega(r, m, carry=d1)
    assert np.array_equal(G2, H2) and c2 == d2 and c2[0] == 0
    # the carried draw starts with the cached variate, then the stream
    assert G2[0, 0] == c1[1]
    assert np.array_equal(G2.ravel(order="F")[1:], G1.ravel(order="F")[:-1])
    hqrrp_amd.set_omega_carry(*c1)
    assert hqrrp_amd.omega_carry() == c1
    hqrrp_amd.set_omega_carry()
    assert hqrrp_amd.omega_carry() == (0, 0.0)
