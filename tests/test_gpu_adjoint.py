"""Q formation, QR without pivoting, and the QR at fixed pivots with its
adjoint, all through the C ABI.

No reference counterpart (the reference's test driver calls LAPACK dorgqr,
WYtest.c:89-93, and dgeqp4 calls dgeqrf for fixed columns, WY.c:221-225), so
these are pinned by independent libraries and by calculus:

* geqrf / orgqr against scipy's LAPACK dgeqrf / dorgqr (same Householder
  convention: the factors agree to rounding), plus Q'Q = I and QR = A;
* the fixed-pivot QR against torch's CPU FP64 linalg.qr of A[:, p];
* the adjoint against torch's CPU FP64 autograd through linalg.qr at the same
  pivots (an independent library gradient), against central finite
  differences (h = 1e-6), and by torch.autograd.gradcheck on 256 x 64.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import hqrrp_amd
from hqrrp_amd import matgen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = np.finfo(np.float64).eps


def _gaussian(m, n, seed):
    return np.random.default_rng(seed).standard_normal((m, n))


@pytest.mark.parametrize("m,n", [(300, 200), (1000, 1000), (700, 130),
                                 (64, 64), (129, 65)])
def test_geqrf_matches_lapack(handle, m, n):
    print(f"geqrf {m}x{n}", flush=True)
    A = _gaussian(m, n, 1)
    dA = hqrrp_amd.to_colmajor(A)
    tau = torch.zeros(min(m, n), dtype=torch.float64, device="cuda")
    hqrrp_amd.geqrf_device(handle, dA, tau, m, n)
    handle.synchronize()
    qr_ref, tau_ref, _, info = lapack.dgeqrf(A)
    assert info == 0
    got = hqrrp_amd.from_colmajor(dA)
    tol = 200 * max(m, n) * EPS
    scale = np.abs(qr_ref).max()
    assert np.abs(got - qr_ref).max() <= tol * scale
    assert np.abs(tau.cpu().numpy() - tau_ref).max() <= tol


@pytest.mark.parametrize("m,n,k", [(300, 200, 200), (1000, 600, 400),
                                   (700, 130, 130), (200, 200, 70)])
def test_orgqr_matches_lapack(handle, m, n, k):
    print(f"orgqr {m}x{n} k={k}", flush=True)
    A = _gaussian(m, n, 2)
    qr_ref, tau_ref, _, _ = lapack.dgeqrf(A)
    Q_ref, _, info = lapack.dorgqr(qr_ref[:, :n], tau_ref[:k])
    assert info == 0
    dA = hqrrp_amd.to_colmajor(qr_ref)
    tau = torch.as_tensor(tau_ref, device="cuda")
    Q = hqrrp_amd.orgqr_device(handle, dA, tau, m, n, k)
    handle.synchronize()
    Qg = hqrrp_amd.from_colmajor(Q)
    assert np.abs(Qg - Q_ref).max() <= 100 * m * EPS
    assert np.abs(Qg.T @ Qg - np.eye(n)).max() <= 100 * m * EPS


def test_orgqr_of_hqrrp_reproduces_ap(handle):
    """Q from the HQRRP reflectors: A P = Q R."""
    m, n = 900, 700
    A = matgen.graded(m, n, 3)
    dA = hqrrp_amd.to_colmajor(A)
    jp = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
    tau = torch.zeros(n, dtype=torch.float64, device="cuda")
    hqrrp_amd.hqrrp_device(handle, dA, jp, tau, m, n)
    Q = hqrrp_amd.orgqr_device(handle, dA, tau, m, n, n)
    handle.synchronize()
    Qg = hqrrp_amd.from_colmajor(Q)
    R = np.triu(hqrrp_amd.from_colmajor(dA)[:n, :])
    AP = A[:, jp.cpu().numpy() - 1]
    assert np.linalg.norm(AP - Qg @ R) <= 50 * n * EPS * np.linalg.norm(A)
    assert np.abs(Qg.T @ Qg - np.eye(n)).max() <= 50 * n * EPS


def _torch_fixed(A, p, k):
    """CPU FP64 reference: (Q, R) of A[:, p-1] truncated to k."""
    AP = A[:, p - 1]
    Q, R = torch.linalg.qr(AP, mode="reduced")
    return Q[:, :k], R[:k, :]


@pytest.mark.parametrize("m,n,k", [(256, 64, 48), (300, 120, 120),
                                   (500, 300, 200), (4096, 1024, 256)])
def test_qr_fixed_matches_torch(handle, m, n, k):
    print(f"qr_fixed {m}x{n} k={k}", flush=True)
    g = torch.Generator().manual_seed(5)
    A = torch.randn(m, n, dtype=torch.float64, generator=g)
    p = torch.randperm(n, generator=g).to(torch.int32) + 1
    Q, R = hqrrp_amd.qr_fixed(handle, A.cuda(), p.cuda(), k)
    handle.synchronize()
    Qr, Rr = _torch_fixed(A, p.long(), k)
    assert (Q.cpu() - Qr).abs().max() <= 1e3 * m * EPS
    assert (R.cpu() - Rr).abs().max() <= 1e3 * m * EPS * Rr.abs().max()


@pytest.mark.parametrize("m,n,k", [(256, 64, 48), (256, 64, 64),
                                   (300, 120, 37), (4096, 1024, 256)])
def test_adjoint_matches_torch_autograd(handle, m, n, k):
    print(f"adjoint {m}x{n} k={k}", flush=True)
    g = torch.Generator().manual_seed(6)
    A = torch.randn(m, n, dtype=torch.float64, generator=g)
    p = torch.randperm(n, generator=g).to(torch.int32) + 1
    gQ = torch.randn(m, k, dtype=torch.float64, generator=g)
    gR = torch.randn(k, n, dtype=torch.float64, generator=g)
    # independent library gradient: torch CPU autograd through linalg.qr
    Ac = A.clone().requires_grad_(True)
    Qr, Rr = _torch_fixed(Ac, p.long(), k)
    ((Qr * gQ).sum() + (Rr * gR).sum()).backward()
    ref = Ac.grad
    Q, R = hqrrp_amd.qr_fixed(handle, A.cuda(), p.cuda(), k)
    gA = hqrrp_amd.qr_fixed_adjoint(handle, A.cuda(), p.cuda(), k, Q, R,
                                    gQ.cuda(), gR.cuda())
    handle.synchronize()
    err = (gA.cpu() - ref).abs().max() / ref.abs().max()
    assert err <= 1e-9, err


def test_adjoint_central_differences(handle):
    """<gA, E> against (f(A+hE) - f(A-hE)) / 2h, f = <gQ, Q> + <gR, R>."""
    m, n, k, h = 256, 64, 48, 1e-6
    g = torch.Generator().manual_seed(7)
    A = torch.randn(m, n, dtype=torch.float64, generator=g).cuda()
    p = (torch.randperm(n, generator=g).to(torch.int32) + 1).cuda()
    gQ = torch.randn(m, k, dtype=torch.float64, generator=g).cuda()
    gR = torch.randn(k, n, dtype=torch.float64, generator=g).cuda()

    def f(X):
        Q, R = hqrrp_amd.qr_fixed(handle, X, p, k)
        handle.synchronize()
        return float((Q * gQ).sum() + (R * gR).sum())

    Q, R = hqrrp_amd.qr_fixed(handle, A, p, k)
    gA = hqrrp_amd.qr_fixed_adjoint(handle, A, p, k, Q, R, gQ, gR)
    handle.synchronize()
    for trial in range(4):
        E = torch.randn(m, n, dtype=torch.float64, generator=g).cuda()
        fd = (f(A + h * E) - f(A - h * E)) / (2 * h)
        an = float((gA * E).sum())
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (trial, fd, an)


def test_autograd_function_gradcheck():
    """torch.autograd.gradcheck of fixed_pivot_qr on 256 x 64 (h = 1e-6)."""
    m, n, k = 256, 64, 48
    g = torch.Generator().manual_seed(8)
    A = torch.randn(m, n, dtype=torch.float64, generator=g).cuda()
    A.requires_grad_(True)
    p = (torch.randperm(n, generator=g).to(torch.int32) + 1).cuda()

    def fn(X):
        return hqrrp_amd.fixed_pivot_qr(X, p, k)

    assert torch.autograd.gradcheck(fn, (A,), eps=1e-6, atol=1e-5, rtol=1e-4,
                                    fast_mode=True)
