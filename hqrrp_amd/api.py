"""Reference-shaped API over the C ABI.

Host entry points take numpy arrays (column-major, float64) exactly like the
reference's C functions take host buffers:

* :func:`dgeqp4` -- LAPACK ``dgeqp3``-compatible driver
  (reference ``NoFLA_HQRRP_WY_blk_var4.c:127-300``);
* :func:`hqrrp` -- ``NoFLA_HQRRP_WY_blk_var4(m, n, A, lda, jpvt, tau, nb_alg,
  pp, panel_pivoting)`` (reference ``NoFLA_HQRRP_WY_blk_var4.c:303-571``).

Device entry points take torch tensors on the GPU.  A column-major m x n
matrix is held as a contiguous torch tensor of shape (n, m) (its transpose in
row-major order), see :func:`to_colmajor`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Options, Trace, check, lib

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def reference_omega(r: int, m: int, seed: int = 12, carry=None):
    """The reference's Gaussian sketch matrix (r x m, column-major) as the
    first call of a process draws it.  With carry=(have, value), the
    generator starts from that cached second variate and the result is
    (G, carry after the draw)."""
    G = np.zeros((r, m), dtype=np.float64, order="F")
    if carry is None:
        lib().hqrrp_reference_omega(r, m, _ptr(G), seed)
        return G
    have = C.c_int(int(carry[0]))
    val = C.c_double(float(carry[1]))
    lib().hqrrp_reference_omega_carry(r, m, _ptr(G), seed, C.byref(have),
                                      C.byref(val))
    return G, (have.value, val.value)


def omega_carry():
    """The library's process-wide cached variate of the reference generator
    (have, value), read and advanced by every factorisation that draws the
    reference Omega."""
    have = C.c_int(0)
    val = C.c_double(0.0)
    lib().hqrrp_omega_carry_get(C.byref(have), C.byref(val))
    return have.value, val.value


def set_omega_carry(have=0, value=0.0):
    """Replace that state ((0, 0.0): as in a new process)."""
    lib().hqrrp_omega_carry_set(int(have), float(value))


def dgeqp4(A, jpvt=None, lwork=None):
    """QR with column pivoting of a copy of A (dgeqp3 calling sequence).

    Returns (A_factored, jpvt (1-based), tau, info, work[0]).
    """
    A = np.asfortranarray(np.array(A, dtype=np.float64, copy=True))
    m, n = A.shape
    jp = (np.zeros(n, dtype=np.int32) if jpvt is None
          else np.array(jpvt, dtype=np.int32, copy=True))
    tau = np.zeros(max(1, min(m, n)))
    lw = max(1, 128 * n) if lwork is None else int(lwork)
    work = np.zeros(max(1, lw))
    info = C.c_int(0)
    lib().dgeqp4(C.byref(C.c_int(m)), C.byref(C.c_int(n)), _ptr(A),
                 C.byref(C.c_int(max(1, m))), _ptr(jp), _ptr(tau), _ptr(work),
                 C.byref(C.c_int(lw)), C.byref(info))
    return A, jp, tau[:min(m, n)], info.value, work[0]


def hqrrp(A, nb_alg=64, pp=10, panel_pivoting=1, jpvt=None):
    """NoFLA_HQRRP_WY_blk_var4 on a copy of A; returns (A, jpvt, tau)."""
    A = np.asfortranarray(np.array(A, dtype=np.float64, copy=True))
    m, n = A.shape
    jp = (np.arange(1, n + 1, dtype=np.int32) if jpvt is None
          else np.array(jpvt, dtype=np.int32, copy=True))
    tau = np.zeros(max(1, min(m, n)))
    rc = lib().NoFLA_HQRRP_WY_blk_var4(m, n, _ptr(A), max(1, m), _ptr(jp),
                                       _ptr(tau), nb_alg, pp, panel_pivoting)
    check(rc, "NoFLA_HQRRP_WY_blk_var4")
    return A, jp, tau[:min(m, n)]


def hqrrp_ut(A, nb_alg=64, pp=10, panel_pivoting=1, p=None):
    """The libflame UT-variant front end (FLA_HQRRP_UT_blk_var2) on a copy
    of A; returns (A, p, s) in libflame's UT convention."""
    A = np.asfortranarray(np.array(A, dtype=np.float64, copy=True))
    m, n = A.shape
    jp = (np.arange(1, n + 1, dtype=np.int32) if p is None
          else np.array(p, dtype=np.int32, copy=True))
    s = np.zeros(max(1, min(m, n)))
    rc = lib().hqrrp_ut_blk_var2(m, n, _ptr(A), max(1, m), _ptr(jp), _ptr(s),
                                 nb_alg, pp, panel_pivoting)
    check(rc, "hqrrp_ut_blk_var2")
    return A, jp, s[:min(m, n)]


# ------------------------------------------------------------ device side
def to_colmajor(A, device="cuda"):
    """numpy/torch m x n -> contiguous torch (n, m) = column-major m x n."""
    import torch

    t = torch.as_tensor(np.asarray(A, dtype=np.float64) if not isinstance(
        A, torch.Tensor) else A, dtype=torch.float64)
    return t.t().contiguous().to(device)


def from_colmajor(T):
    """contiguous torch (n, m) -> numpy m x n (Fortran order)."""
    return np.asfortranarray(T.detach().cpu().numpy().T)


class Handle:
    """Owns a library handle (device workspace, barrier words, a stream)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().hqrrp_create(C.byref(self._h), device), "hqrrp_create")
        self.device = device

    @property
    def ptr(self):
        return self._h

    def use_stream(self, stream) -> None:
        """Run this handle's work on a torch.cuda.Stream, ordered with torch's
        work on it (torch's default stream is the null stream); None: the
        handle's own stream."""
        if stream is None:
            check(lib().hqrrp_reset_stream(self._h), "hqrrp_reset_stream")
        else:
            check(lib().hqrrp_set_stream(self._h,
                                         C.c_void_p(stream.cuda_stream)),
                  "hqrrp_set_stream")

    def set_panel_alg(self, alg: int) -> None:
        """0 (default): Cholesky-QR panel with the per-step kernel as its
        fallback; 1: the per-step panel kernel for every panel."""
        check(lib().hqrrp_set_panel_alg(self._h, int(alg)),
              "hqrrp_set_panel_alg")

    def synchronize(self) -> None:
        check(lib().hqrrp_synchronize(self._h), "hqrrp_synchronize")

    def profile(self, on: bool = True, phases=None) -> None:
        """Start (and reset) or stop per-phase event timing; `phases` (names)
        limits the recording to those phases (default: all)."""
        L = lib()
        mask = 0xFFFFFFFF
        if phases is not None:
            names = [L.hqrrp_phase_name(i).decode()
                     for i in range(L.hqrrp_phase_count())]
            mask = 0
            for p in phases:
                mask |= 1 << names.index(p)
        check(L.hqrrp_profile_phases(self._h, mask), "hqrrp_profile_phases")
        check(L.hqrrp_profile_enable(self._h, 1 if on else 0),
              "hqrrp_profile_enable")

    def profile_read(self) -> dict:
        """{phase: (ms, launches, algorithmic flops, algorithmic bytes)}."""
        L = lib()
        n = L.hqrrp_phase_count()
        ms = (C.c_double * n)()
        la = (C.c_longlong * n)()
        fl = (C.c_double * n)()
        by = (C.c_double * n)()
        check(L.hqrrp_profile_read(self._h, n, ms, la, fl, by),
              "hqrrp_profile_read")
        return {L.hqrrp_phase_name(i).decode(): (ms[i], la[i], fl[i], by[i])
                for i in range(n)}

    def batch_tune(self, pp: int = 8, panel_pivoting: int = 1) -> np.ndarray:
        """Autotune the batched path's block size per 64 x 64 shape bucket;
        returns the 8 x 8 table ([m bucket, n bucket])."""
        check(lib().hqrrp_batch_tune(self._h, pp, panel_pivoting),
              "hqrrp_batch_tune")
        t = np.zeros(64, dtype=np.int32)
        check(lib().hqrrp_batch_tune_table(self._h, _ptr(t)),
              "hqrrp_batch_tune_table")
        return t.reshape(8, 8)

    def close(self) -> None:
        if self._h:
            lib().hqrrp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def options(nb=64, pp=10, panel_pivoting=1, seed=12, omega=None,
            max_rank=0, sketch_prec=0, lookahead=None, trace=None) -> Options:
    """Options struct; `trace` is a dict of device tensors with any of the
    keys of ``hqrrp_trace`` (sketch_cols, y0, y, g) to read back."""
    o = Options()
    lib().hqrrp_default_options(C.byref(o))
    o.nb, o.pp, o.panel_pivoting, o.seed = nb, pp, panel_pivoting, seed
    o.max_rank = max_rank
    o.sketch_prec = sketch_prec
    if lookahead is not None:
        o.lookahead = int(lookahead)
    if omega is not None:
        o.omega_kind = 1
        o.omega = C.c_void_p(omega.data_ptr())
    if trace:
        t = Trace()
        for k, v in trace.items():
            setattr(t, k, C.c_void_p(v.data_ptr()))
        o._trace_keep = t  # the struct must outlive the call
        o.trace = C.pointer(t)
    return o


def hqrrp_device(handle: Handle, At, jpvt_t, tau_t, m: int, n: int,
                 lda: int | None = None, **kw) -> None:
    """Factor the column-major m x n matrix held by torch tensor At in place.

    jpvt_t: int32 tensor (n), permuted in place; tau_t: float64 (min(m,n)).
    Stream-ordered on the handle's stream; call handle.synchronize().
    """
    o = options(**kw)
    rc = lib().hqrrp_factor_dev(handle.ptr, m, n, C.c_void_p(At.data_ptr()),
                                m if lda is None else lda,
                                C.c_void_p(jpvt_t.data_ptr()),
                                C.c_void_p(tau_t.data_ptr()), C.byref(o))
    check(rc, "hqrrp_factor_dev")


def hqrrp_ut_device(handle: Handle, At, p_t, s_t, m: int, n: int,
                    lda: int | None = None, **kw) -> None:
    """hqrrp_ut_factor_dev: hqrrp_device in libflame's UT convention (s_t
    receives s = 1/tau, 1/2 where LAPACK's tau is 0, with R's row negated
    there)."""
    o = options(**kw)
    rc = lib().hqrrp_ut_factor_dev(handle.ptr, m, n, C.c_void_p(At.data_ptr()),
                                   m if lda is None else lda,
                                   C.c_void_p(p_t.data_ptr()),
                                   C.c_void_p(s_t.data_ptr()), C.byref(o))
    check(rc, "hqrrp_ut_factor_dev")


# ------------------------------------- Q and the fixed-pivot QR with adjoint
def _vp(t):
    return C.c_void_p(t.data_ptr())


def geqrf_device(handle: Handle, At, tau_t, m: int, n: int,
                 lda: int | None = None) -> None:
    """Householder QR without pivoting of the column-major m x n matrix held
    by At (shape (n, lda)), in place; tau_t: float64 (min(m,n))."""
    check(lib().hqrrp_geqrf_dev(handle.ptr, m, n, _vp(At),
                                m if lda is None else lda, _vp(tau_t)),
          "hqrrp_geqrf_dev")


def geqp3_device(handle: Handle, At, jpvt_t, tau_t, m: int, n: int,
                 lda: int | None = None) -> None:
    """Classic greedy QRCP (LAPACK dgeqp3 semantics) of the column-major
    m x n matrix held by At, in place; jpvt_t (int32, n) permuted in place,
    tau_t float64 (min(m,n)).  Stream-ordered on the handle's stream."""
    check(lib().hqrrp_geqp3_dev(handle.ptr, m, n, _vp(At),
                                m if lda is None else lda, _vp(jpvt_t),
                                _vp(tau_t)), "hqrrp_geqp3_dev")


def orgqr_device(handle: Handle, At, tau_t, m: int, n: int, k: int,
                 lda: int | None = None):
    """Q (column-major m x n, returned as a torch (n, m) tensor) from the
    first k reflectors of At (dorgqr semantics)."""
    import torch

    Q = torch.empty((n, m), dtype=torch.float64, device=At.device)
    check(lib().hqrrp_orgqr_dev(handle.ptr, m, n, k, _vp(At),
                                m if lda is None else lda, _vp(tau_t), _vp(Q),
                                m), "hqrrp_orgqr_dev")
    return Q


def ormqr_device(handle: Handle, side: str, trans: str, At, tau_t, k: int,
                 Ct, m: int, n: int, lda: int | None = None) -> None:
    """op(Q) C or C op(Q) in place (LAPACK dormqr semantics) on the
    column-major m x n matrix held by Ct, with Q from the first k reflectors
    of At (column-major, ld lda: m for side 'L', n for 'R')."""
    nq = m if side.upper() == "L" else n
    check(lib().hqrrp_ormqr_dev(handle.ptr, side.encode(), trans.encode(), m,
                                n, k, _vp(At), nq if lda is None else lda,
                                _vp(tau_t), _vp(Ct), m), "hqrrp_ormqr_dev")


def _colmajor_of(A):
    """torch m x n (any layout) -> contiguous (n, m): column-major m x n."""
    return A.detach().t().contiguous()


def qr_fixed(handle: Handle, A, jpvt, k: int):
    """QR of A[:, jpvt-1] truncated to k: returns Q (m x k) and R (k x n) as
    ordinary torch tensors on A's device (jpvt: 1-based int32 on the device).
    """
    import torch

    m, n = A.shape
    Acm = _colmajor_of(A.to(torch.float64))
    Qcm = torch.empty((k, m), dtype=torch.float64, device=A.device)
    Rcm = torch.empty((n, k), dtype=torch.float64, device=A.device)
    jp = jpvt.to(device=A.device, dtype=torch.int32).contiguous()
    check(lib().hqrrp_qr_fixed_dev(handle.ptr, m, n, k, _vp(Acm), m, _vp(jp),
                                   _vp(Qcm), m, _vp(Rcm), k),
          "hqrrp_qr_fixed_dev")
    return Qcm.t(), Rcm.t()


def qr_fixed_adjoint(handle: Handle, A, jpvt, k: int, Q, R, gQ, gR):
    """Cotangent of A for qr_fixed at fixed pivots (torch m x n)."""
    import torch

    m, n = A.shape
    f64 = torch.float64
    Acm = _colmajor_of(A.to(f64))
    Qcm, Rcm = _colmajor_of(Q.to(f64)), _colmajor_of(R.to(f64))
    z = lambda shape: torch.zeros(shape, dtype=f64, device=A.device)
    gQcm = _colmajor_of(gQ.to(f64)) if gQ is not None else z((k, m))
    gRcm = _colmajor_of(gR.to(f64)) if gR is not None else z((n, k))
    gAcm = torch.empty((n, m), dtype=f64, device=A.device)
    jp = jpvt.to(device=A.device, dtype=torch.int32).contiguous()
    check(lib().hqrrp_qr_fixed_adjoint_dev(
        handle.ptr, m, n, k, _vp(Acm), m, _vp(jp), _vp(Qcm), m, _vp(Rcm), k,
        _vp(gQcm), m, _vp(gRcm), k, _vp(gAcm), m),
        "hqrrp_qr_fixed_adjoint_dev")
    return gAcm.t()


_DEFAULT_HANDLES = {}


def _default_handle(device) -> Handle:
    import torch

    idx = torch.device(device).index or 0
    if idx not in _DEFAULT_HANDLES:
        h = Handle(idx)
        _DEFAULT_HANDLES[idx] = h
    h = _DEFAULT_HANDLES[idx]
    h.use_stream(torch.cuda.current_stream(idx))
    return h


def _fixed_qr_function():
    import torch

    class FixedPivotQR(torch.autograd.Function):
        """(Q, R) = QR of A[:, jpvt-1] truncated to k columns of Q, with the
        pivots held fixed; backward through the library's adjoint."""

        @staticmethod
        def forward(ctx, A, jpvt, k):
            h = _default_handle(A.device)
            Q, R = qr_fixed(h, A, jpvt, k)
            ctx.save_for_backward(A, jpvt, Q, R)
            ctx.k = k
            return Q, R

        @staticmethod
        def backward(ctx, gQ, gR):
            A, jpvt, Q, R = ctx.saved_tensors
            h = _default_handle(A.device)
            gA = qr_fixed_adjoint(h, A, jpvt, ctx.k, Q, R, gQ, gR)
            return gA.to(A.dtype), None, None

    return FixedPivotQR


def fixed_pivot_qr(A, jpvt, k: int):
    """Differentiable (Q, R) of A[:, jpvt-1] truncated to k (see
    hqrrp_qr_fixed_dev); gradients by hqrrp_qr_fixed_adjoint_dev."""
    return _fixed_qr_function().apply(A, jpvt, k)


# ------------------------------------------------------------ batched path
class BatchPlan:
    """A prepared batch of small matrices (m, n <= 512) for the batched
    persistent launch: shapes, per-matrix block size (nb: None = chosen per
    shape, an int for all, or one per matrix), oversampling pp, queue order
    and workspace.  See hqrrp_batch_create in include/hqrrp_amd.h."""

    def __init__(self, handle: Handle, m, n, nb=None, pp: int = 8,
                 panel_pivoting: int = 1, lda=None):
        m = np.ascontiguousarray(m, dtype=np.int32)
        n = np.ascontiguousarray(n, dtype=np.int32)
        if m.shape != n.shape:
            raise ValueError("m and n must have the same length")
        self.count = int(m.size)
        self.m, self.n = m, n
        nb_a = None
        if nb is not None:
            nb_a = np.ascontiguousarray(
                np.full(self.count, nb) if np.isscalar(nb) else nb,
                dtype=np.int32)
        lda_a = None if lda is None else np.ascontiguousarray(lda, np.int32)
        self._p = C.c_void_p()
        self._keep = (m, n, nb_a, lda_a)
        check(lib().hqrrp_batch_create(
            handle.ptr, self.count, _ptr(m), _ptr(n),
            None if lda_a is None else _ptr(lda_a),
            None if nb_a is None else _ptr(nb_a), pp, panel_pivoting,
            C.byref(self._p)), "hqrrp_batch_create")
        self.pp, self.panel_pivoting = pp, panel_pivoting

    @property
    def nb(self) -> np.ndarray:
        out = np.zeros(max(1, self.count), dtype=np.int32)
        check(lib().hqrrp_batch_get_nb(self._p, _ptr(out)),
              "hqrrp_batch_get_nb")
        return out[:self.count]

    def factor(self, handle: Handle, A_ptrs, jpvt_ptrs, tau_ptrs) -> None:
        """Factor in place; the arguments are int64 torch tensors on the
        device holding the matrices', pivot vectors' and tau vectors'
        device addresses."""
        check(lib().hqrrp_batch_factor_dev(handle.ptr, self._p,
                                           _vp(A_ptrs), _vp(jpvt_ptrs),
                                           _vp(tau_ptrs)),
              "hqrrp_batch_factor_dev")

    def profile(self, on: bool = True) -> None:
        """Start (and reset) or stop the kernel's per-phase cycle counts."""
        check(lib().hqrrp_batch_profile(self._p, 1 if on else 0),
              "hqrrp_batch_profile")

    def profile_read(self) -> dict:
        """{phase: shader-clock cycles summed over workgroups}."""
        L = lib()
        k = L.hqrrp_batch_phase_count()
        out = (C.c_double * k)()
        check(L.hqrrp_batch_profile_read(self._p, k, out),
              "hqrrp_batch_profile_read")
        return {L.hqrrp_batch_phase_name(i).decode(): out[i]
                for i in range(k)}

    def close(self) -> None:
        if self._p:
            lib().hqrrp_batch_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PackedBatch:
    """A batch packed in three device buffers (matrices column-major with
    ld = m, pivots initialised to 1..n, tau) and the device pointer arrays
    BatchPlan.factor takes."""

    def __init__(self, m, n, device="cuda"):
        import torch

        self.m = np.asarray(m, dtype=np.int64)
        self.n = np.asarray(n, dtype=np.int64)
        sizes = self.m * self.n
        self.a_off = np.concatenate([[0], np.cumsum(sizes)])
        self.j_off = np.concatenate([[0], np.cumsum(self.n)])
        k = np.minimum(self.m, self.n)
        self.t_off = np.concatenate([[0], np.cumsum(np.maximum(k, 1))])
        self.A = torch.zeros(int(max(1, self.a_off[-1])), dtype=torch.float64,
                             device=device)
        self.jpvt = torch.zeros(int(max(1, self.j_off[-1])), dtype=torch.int32,
                                device=device)
        self.tau = torch.zeros(int(max(1, self.t_off[-1])),
                               dtype=torch.float64, device=device)
        self.reset_pivots()
        base = lambda t, off, es: torch.as_tensor(  # noqa: E731
            t.data_ptr() + off[:-1] * es, dtype=torch.int64).to(device)
        self.A_ptrs = base(self.A, self.a_off, 8)
        self.jpvt_ptrs = base(self.jpvt, self.j_off, 4)
        self.tau_ptrs = base(self.tau, self.t_off, 8)

    def reset_pivots(self) -> None:
        import torch

        host = np.concatenate([np.arange(1, int(c) + 1, dtype=np.int32)
                               for c in self.n]) if self.n.size else \
            np.zeros(0, np.int32)
        if host.size:
            self.jpvt.copy_(torch.as_tensor(host))

    def load(self, mats) -> None:
        """Copy host matrices (numpy, shape m_i x n_i) into the buffer."""
        import torch

        flat = np.concatenate([np.asarray(a, dtype=np.float64).ravel(order="F")
                               for a in mats]) if mats else np.zeros(0)
        if flat.size:
            self.A[:flat.size].copy_(torch.as_tensor(flat))

    def unpack(self):
        """[(A_i factored, jpvt_i, tau_i)] as numpy arrays."""
        A = self.A.cpu().numpy()
        J = self.jpvt.cpu().numpy()
        T = self.tau.cpu().numpy()
        out = []
        for i in range(self.m.size):
            m, n = int(self.m[i]), int(self.n[i])
            a = A[self.a_off[i]:self.a_off[i + 1]].reshape((m, n), order="F")
            out.append((np.asfortranarray(a),
                        J[self.j_off[i]:self.j_off[i + 1]].copy(),
                        T[self.t_off[i]:self.t_off[i] + min(m, n)].copy()))
        return out


def hqrrp_batched(mats, nb=None, pp: int = 8, panel_pivoting: int = 1,
                  handle: Handle | None = None):
    """NoFLA_HQRRP_WY_blk_var4 on copies of many small host matrices in one
    batched launch; returns [(A, jpvt, tau)] like :func:`hqrrp`."""
    h = handle if handle is not None else _default_handle(0)
    m = [np.shape(a)[0] for a in mats]
    n = [np.shape(a)[1] for a in mats]
    pk = PackedBatch(m, n)
    pk.load(mats)
    plan = BatchPlan(h, m, n, nb=nb, pp=pp, panel_pivoting=panel_pivoting)
    plan.factor(h, pk.A_ptrs, pk.jpvt_ptrs, pk.tau_ptrs)
    h.synchronize()
    out = pk.unpack()
    plan.close()
    return out
