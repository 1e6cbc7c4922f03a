"""The full-size factorizations (16384 x 16384 and ragged rectangles around
10^4), checked through size-independent properties (the headline's pivots
are compared with the oracle itself in tests/test_gpu_headline.py):

* jpvt is a permutation of 1..n;
* ||AP - QR||_F / ||A||_F and max|Q'Q - I| <= 20 sqrt(max(m, n)) eps (the
  growth of Householder QR in practice; 5.7e-13 at 16384, where 4.7e-15 and
  1.2e-14 are measured), with Q formed from the reflectors by the library's
  own dorgqr;
* for A = U diag(sigma) V' with sigma_k = 10^(-12 k / n), |R_kk| stays
  within three decades of sigma_k while sigma_k >= 1e-6 (the diagonal of a
  rank-revealing pivoted QR tracks the spectrum; measured [0.136, 1.1]);
* the look-ahead schedule is bitwise equal to the sequential one.

Everything stays on the device; the products use torch's FP64 GEMM.
"""
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = 2.0 ** -52


def _factor(handle, At, m, n, **kw):
    import hqrrp_amd

    jp = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
    tau = torch.zeros(max(1, min(m, n)), dtype=torch.float64, device="cuda")
    hqrrp_amd.hqrrp_device(handle, At, jp, tau, m, n, **kw)
    handle.synchronize()
    torch.cuda.synchronize()
    return jp, tau


def _check(handle, A0t, At, jp, tau, m, n):
    """A0t/At: (n, m) tensors holding the column-major input / output."""
    import hqrrp_amd

    k = min(m, n)
    assert torch.equal(torch.sort(jp.long()).values,
                       torch.arange(1, n + 1, device="cuda"))
    Qt = hqrrp_amd.orgqr_device(handle, At.clone(), tau, m, k, k)
    handle.synchronize()
    torch.cuda.synchronize()
    Q = Qt.t()                                  # m x k
    R = torch.triu(At.t()[:k, :])               # k x n
    AP = A0t[(jp.long() - 1)].t()               # m x n
    res = (torch.linalg.matrix_norm(AP - Q @ R)
           / torch.linalg.matrix_norm(AP)).item()
    orth = (Qt @ Q - torch.eye(k, dtype=torch.float64, device="cuda")
            ).abs().max().item()
    tol = 20 * max(m, n) ** 0.5 * EPS
    print(f"{m}x{n}: residual {res:.2e} orthogonality {orth:.2e} "
          f"(bound {tol:.1e})", flush=True)
    assert res <= tol and orth <= tol
    return R


@pytest.mark.parametrize("m,n", [(16384, 16384)])
def test_full_size_graded(handle, m, n):
    from hqrrp_amd import matgen

    print(f"full size graded {m}x{n}", flush=True)
    A0t = matgen.graded_device(m, n, seed=3)
    At = A0t.clone()
    jp, tau = _factor(handle, At, m, n)
    R = _check(handle, A0t, At, jp, tau, m, n)
    k = min(m, n)
    sig = 10.0 ** (-12.0 * torch.arange(k, device="cuda",
                                        dtype=torch.float64) / k)
    live = sig >= 1e-6
    ratio = (torch.diagonal(R).abs()[live] / sig[live])
    print(f"|R_kk| / sigma_k in [{ratio.min().item():.3g}, "
          f"{ratio.max().item():.3g}] over {int(live.sum())} columns",
          flush=True)
    assert ratio.min().item() >= 1e-3 and ratio.max().item() <= 1e3


@pytest.mark.parametrize("m,n,family", [(12001, 9000, "lowrank"),
                                        (9000, 12001, "gaussian")])
def test_ragged_rectangles(handle, m, n, family):
    from hqrrp_amd import matgen

    print(f"ragged {family} {m}x{n}", flush=True)
    if family == "lowrank":
        A0t = matgen.lowrank_device(m, n, rank=700, seed=4)
    else:
        A0t = matgen.gaussian_device(m, n, seed=5)
    At = A0t.clone()
    jp, tau = _factor(handle, At, m, n)
    _check(handle, A0t, At, jp, tau, m, n)


def test_full_size_lookahead_bitwise(handle):
    """At 16384 x 16384 the overlapped schedule matches the sequential one
    bit for bit (pivots, tau and every entry of the factored matrix)."""
    from hqrrp_amd import matgen

    m = n = 16384
    print(f"full size lookahead {m}x{n}", flush=True)
    A0t = matgen.graded_device(m, n, seed=6)
    A1t = A0t.clone()
    j0, t0 = _factor(handle, A0t, m, n, lookahead=0)
    j1, t1 = _factor(handle, A1t, m, n, lookahead=1)
    assert torch.equal(j0, j1)
    assert torch.equal(t0, t1)
    assert torch.equal(A0t, A1t)
