"""CPU oracle for the HQRRP hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It wraps two shared objects:

* ``liboracle.so`` -- the C restatement of the reference driver
  (``lapack_compatible_sources/NoFLA_HQRRP_WY_blk_var4.c``), which issues the
  same LAPACK calls in the same order; LAPACK is scipy's bundled OpenBLAS,
  bound at run time;
* ``_ref/libhqrrp_ref.so`` -- the reference source itself compiled against the
  same OpenBLAS (optional: built only where the reference tree is present).
"""
from __future__ import annotations

import ctypes as C
import glob
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)


def openblas_path() -> str:
    import scipy

    libs = os.path.dirname(scipy.__file__) + ".libs"
    cands = sorted(glob.glob(os.path.join(libs, "libscipy_openblas*.so")))
    if not cands:
        raise RuntimeError("scipy's bundled OpenBLAS not found")
    return cands[0]


def lib():
    """Load liboracle.so and bind LAPACK (builds nothing)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        rc = L.orc_init(openblas_path().encode())
        if rc != 0:
            raise RuntimeError(f"orc_init failed ({rc})")
        _LIB = L
    return _LIB


def ref():
    """The compiled reference, or None where it was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libhqrrp_ref.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path)
    return _REF


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def _dp(a):
    return a.ctypes.data_as(_d)


def _ip(a):
    return a.ctypes.data_as(_i)


def fortran(a, dtype=np.float64):
    return np.asfortranarray(np.array(a, dtype=dtype, copy=True))


# ---------------------------------------------------------------- generators
class RandStream:
    """Reentrant restatement of glibc ``rand()`` (seeded like ``srand``)."""

    def __init__(self, seed: int):
        self._st = (C.c_uint32 * 33)()
        lib().orc_rand_seed(self._st, C.c_uint(seed))

    def next(self) -> int:
        return int(lib().orc_rand_next(self._st))


def omega(r: int, m: int, seed: int = 12, carry=None):
    """The reference's Gaussian sketch matrix G (r x m, column-major), as the
    first call of a process draws it; with carry=(have, value) -- the
    generator's statics as a later call finds them -- returns (G, the
    statics the draw leaves)."""
    G = np.zeros((r, m), dtype=np.float64, order="F")
    if carry is None:
        lib().orc_omega(C.c_int(r), C.c_int(m), _dp(G), C.c_int(r),
                        C.c_uint(seed))
        return G
    h = C.c_int(int(carry[0]))
    v = C.c_double(float(carry[1]))
    lib().orc_omega_carry(C.c_int(r), C.c_int(m), _dp(G), C.c_int(r),
                          C.c_uint(seed), C.byref(h), C.byref(v))
    return G, (h.value, v.value)


def ref_omega(r: int, m: int) -> np.ndarray:
    R = ref()
    G = np.zeros((r, m), dtype=np.float64, order="F")
    R.ref_omega(C.c_int(r), C.c_int(m), _dp(G), C.c_int(r))
    return G


# ---------------------------------------------------------------- drivers
def hqrrp(A, nb=64, pp=10, panel_pivoting=1, jpvt=None, omega_in=None,
          max_blocks=0, trace=False, use_ref=False, max_cols=0):
    """Factor a copy of A; returns (A_out, jpvt (1-based), tau[, traces]).

    trace=True appends (sketch gaps, panel gaps, sketch order): per pivot
    position, the (largest, second largest) candidate partial norms of the
    sketch and of the panel decision, and the columns each block's sketch
    chose in its own order (jpvt values).  max_cols > 0 truncates (see
    orc_hqrrp_ex)."""
    A = fortran(A)
    m, n = A.shape
    jp = (np.arange(1, n + 1, dtype=np.int32) if jpvt is None
          else np.array(jpvt, dtype=np.int32))
    tau = np.zeros(max(min(m, n), 1))
    if use_ref:
        R = ref()
        if R is None:
            raise RuntimeError("reference not built")
        R.NoFLA_HQRRP_WY_blk_var4(C.c_int(m), C.c_int(n), _dp(A), C.c_int(m),
                                  _ip(jp), _dp(tau), C.c_int(nb), C.c_int(pp),
                                  C.c_int(panel_pivoting))
        return A, jp, tau[:min(m, n)]
    mn = min(m, n)
    ts = np.zeros(2 * max(mn, 1)) if trace else None
    tp = np.zeros(2 * max(mn, 1)) if trace else None
    sc = np.zeros(max(mn, 1), dtype=np.int32) if trace else None
    om = None if omega_in is None else fortran(omega_in)
    lib().orc_hqrrp_ex(
        C.c_int(m), C.c_int(n), _dp(A), C.c_int(m), _ip(jp), _dp(tau),
        C.c_int(nb), C.c_int(pp), C.c_int(panel_pivoting),
        _dp(om) if om is not None else None, C.c_int(max_blocks),
        _dp(ts) if trace else None, _dp(tp) if trace else None, None, None,
        C.c_int(max_cols), _ip(sc) if trace else None)
    if trace:
        return (A, jp, tau[:mn], ts.reshape(-1, 2)[:mn],
                tp.reshape(-1, 2)[:mn], sc[:mn])
    return A, jp, tau[:mn]


def hqrrp_ut(A, nb=64, pp=10, panel_pivoting=1, p=None, carry=None):
    """The libflame UT variant (FLA_HQRRP_UT_blk_var2) restated; returns
    (A_out, p (1-based), s) in libflame's UT convention
    (H_k = I - u_k u_k' / s_k)."""
    A = fortran(A)
    m, n = A.shape
    jp = (np.arange(1, n + 1, dtype=np.int32) if p is None
          else np.array(p, dtype=np.int32))
    sv = np.zeros(max(min(m, n), 1))
    h = C.c_int(int(carry[0]) if carry else 0)
    v = C.c_double(float(carry[1]) if carry else 0.0)
    lib().orc_hqrrp_ut(C.c_int(m), C.c_int(n), _dp(A), C.c_int(max(1, m)),
                       _ip(jp), _dp(sv), C.c_int(nb), C.c_int(pp),
                       C.c_int(panel_pivoting), C.byref(h), C.byref(v))
    return A, jp, sv[:min(m, n)]


def ut_q(Af, s):
    """Q = I - U inv(T) U' of the UT transform (U unit lower trapezoidal
    from Af, T upper triangular with diag(T) = s and T(i, j) = u_i'u_j)."""
    m, n = Af.shape
    k = len(s)
    U = np.tril(Af[:, :k], -1)
    U[np.arange(k), np.arange(k)] = 1.0
    T = np.triu(U.T @ U, 1)
    T[np.arange(k), np.arange(k)] = s
    return np.eye(m) - U @ np.linalg.solve(T, U.T)


def dgeqp4(A, jpvt=None, lwork=None, use_ref=False):
    """The dgeqp3-compatible front end; returns (A, jpvt, tau, info, work0)."""
    A = fortran(A)
    m, n = A.shape
    jp = (np.zeros(n, dtype=np.int32) if jpvt is None
          else np.array(jpvt, dtype=np.int32))
    tau = np.zeros(max(min(m, n), 1))
    lw = lwork if lwork is not None else max(1, 128 * n)
    work = np.zeros(max(lw, 1))
    info = C.c_int(0)
    args = (C.byref(C.c_int(m)), C.byref(C.c_int(n)), _dp(A),
            C.byref(C.c_int(max(1, m))), _ip(jp), _dp(tau), _dp(work),
            C.byref(C.c_int(lw)), C.byref(info))
    if use_ref:
        ref().dgeqp4(*args)
    else:
        lib().orc_dgeqp4(*args)
    return A, jp, tau[:min(m, n)], info.value, work[0]


def dgeqp3(A, jpvt=None):
    A = fortran(A)
    m, n = A.shape
    jp = (np.zeros(n, dtype=np.int32) if jpvt is None
          else np.array(jpvt, dtype=np.int32))
    tau = np.zeros(max(min(m, n), 1))
    lw = 3 * n + 1 + 64 * (n + 1)
    work = np.zeros(lw)
    info = C.c_int(0)
    lib().orc_dgeqp3(C.byref(C.c_int(m)), C.byref(C.c_int(n)), _dp(A),
                     C.byref(C.c_int(max(1, m))), _ip(jp), _dp(tau),
                     _dp(work), C.byref(C.c_int(lw)), C.byref(info))
    return A, jp, tau[:min(m, n)], info.value


def form_q(Afac, tau, ncols=None):
    """Explicit Q (m x k) from the Householder vectors below the diagonal."""
    m, n = Afac.shape
    k = len(tau)
    ncols = k if ncols is None else ncols
    Q = np.zeros((m, ncols), order="F")
    Q[:, :min(n, ncols)] = Afac[:, :min(n, ncols)]
    lw = max(1, 64 * ncols)
    work = np.zeros(lw)
    info = C.c_int(0)
    t = np.ascontiguousarray(tau, dtype=np.float64)
    lib().orc_dorgqr(C.byref(C.c_int(m)), C.byref(C.c_int(ncols)),
                     C.byref(C.c_int(k)), _dp(Q), C.byref(C.c_int(m)), _dp(t),
                     _dp(work), C.byref(C.c_int(lw)), C.byref(info))
    assert info.value == 0
    return Q


def qrpmod(A, pivoting=1, num_stages=-1, B=None, Cm=None, build_T=False,
           p=None, trace=False, use_ref=False):
    """The unblocked QRCP helper on copies; returns a dict of outputs."""
    A = fortran(A)
    m, n = A.shape
    p = np.arange(1, n + 1, dtype=np.int32) if p is None else np.array(p, np.int32)
    t = np.zeros(max(n, 1))
    B = fortran(B) if B is not None else np.zeros((1, n), order="F")
    Cm = fortran(Cm) if Cm is not None else np.zeros((1, n), order="F")
    ns = min(m, n) if num_stages < 0 else num_stages
    T = np.zeros((max(ns, 1), max(ns, 1)), order="F")
    tr = np.zeros(2 * max(ns, 1))
    args = [C.c_int(pivoting), C.c_int(num_stages), C.c_int(m), C.c_int(n),
            _dp(A), C.c_int(m), _ip(p), _dp(t),
            C.c_int(1), C.c_int(B.shape[0]), _dp(B), C.c_int(B.shape[0]),
            C.c_int(1), C.c_int(Cm.shape[0]), _dp(Cm), C.c_int(Cm.shape[0]),
            C.c_int(1 if build_T else 0), _dp(T), C.c_int(T.shape[0])]
    if use_ref:
        ref().ref_qrpmod(*args)
    else:
        lib().orc_qrpmod(*args, _dp(tr) if trace else None)
    out = dict(A=A, p=p, tau=t[:ns], B=B, C=Cm, T=T[:ns, :ns] if build_T else None)
    if trace:
        out["trace"] = tr.reshape(-1, 2)[:ns]
    return out


def downdate_y(U, R12, Y2, T, G, use_ref=False):
    """Sketch downdate on copies.

    U: (m-j) x b panel holding the unit-lower reflectors, R12: b x nY2,
    Y2: r x nY2, T: b x b, G: r x (m-j) (= [G1 G2]).  Returns (Y2, G).
    """
    U = fortran(U)
    R12 = fortran(R12)
    Y2 = fortran(Y2)
    T = fortran(T)
    G = fortran(G)
    mU, b = U.shape
    r, nY2 = Y2.shape
    mU21 = mU - b
    nG2 = G.shape[1] - b
    args = [C.c_int(b), C.c_int(mU21), _dp(U), C.c_int(mU), _dp(R12),
            C.c_int(R12.shape[0]), C.c_int(nY2), _dp(Y2), C.c_int(r), _dp(T),
            C.c_int(b), C.c_int(r), _dp(G), C.c_int(nG2),
            G[:, b:].ctypes.data_as(_d) if nG2 > 0 else _dp(G), C.c_int(r)]
    if use_ref:
        ref().ref_downdate_y(*args)
    else:
        lib().orc_downdate_y(*args)
    return Y2, G


def dormqr(side, trans, Afac, tau, Cm):
    """LAPACK dormqr on a copy of Cm: op(Q) C ('L') or C op(Q) ('R'), Q the
    product of the len(tau) reflectors stored below Afac's diagonal."""
    A = fortran(Afac)
    Cm = fortran(Cm)
    m, n = Cm.shape
    k = len(tau)
    lw = max(1, 64 * max(m, n))
    work = np.zeros(lw)
    info = C.c_int(0)
    t = np.ascontiguousarray(tau, dtype=np.float64)
    lib().orc_dormqr(side.encode(), trans.encode(), C.byref(C.c_int(m)),
                     C.byref(C.c_int(n)), C.byref(C.c_int(k)), _dp(A),
                     C.byref(C.c_int(A.shape[0])), _dp(t), _dp(Cm),
                     C.byref(C.c_int(m)), _dp(work), C.byref(C.c_int(lw)),
                     C.byref(info))
    assert info.value == 0, info.value
    return Cm


def dlarft(V, tau):
    V = fortran(V)
    n, k = V.shape
    T = np.zeros((k, k), order="F")
    t = np.ascontiguousarray(tau, dtype=np.float64)
    lib().orc_dlarft(C.c_int(n), C.c_int(k), _dp(V), C.c_int(n), _dp(t), _dp(T),
                     C.c_int(k))
    return T


def apply_q_left_t(U, T, Bm):
    U = fortran(U)
    T = fortran(T)
    Bm = fortran(Bm)
    mU, b = U.shape
    lib().orc_apply_q_left_t(C.c_int(mU), C.c_int(b), _dp(U), C.c_int(mU),
                             _dp(T), C.c_int(T.shape[0]), C.c_int(Bm.shape[0]),
                             C.c_int(Bm.shape[1]), _dp(Bm), C.c_int(Bm.shape[0]))
    return Bm


# ---------------------------------------------------------------- checks
def residuals(A0, Afac, jpvt, tau):
    """(||A P - Q R||_F / ||A||_F, ||Q'Q - I||_F) in FP64 (economy Q)."""
    A0 = np.asarray(A0, dtype=np.float64)
    m, n = A0.shape
    k = min(m, n)
    Q = form_q(Afac, tau)
    R = np.triu(Afac[:k, :])
    AP = A0[:, np.asarray(jpvt) - 1]
    res = np.linalg.norm(AP - Q @ R) / max(np.linalg.norm(A0), 1e-300)
    orth = np.linalg.norm(Q.T @ Q - np.eye(k))
    return float(res), float(orth)
