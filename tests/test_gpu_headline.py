"""The benchmark workloads at their full sizes against the oracle.

* The headline: the 16384 x 16384 graded matrix of the benchmark (Haar U, V,
  sigma_i = 10^(-12 i/n), nb = 64, pp = 10, Omega = the reference stream,
  seed 12) factored on the device and by the oracle (a CPU restatement of
  NoFLA_HQRRP_WY_blk_var4.c:303-571, bitwise equal to the compiled
  reference) from the same input bytes: jpvt equal, or first diverging at a
  near-tie of the exact decision (tests.helpers.check_pivots, which prints
  the divergence); |diag R| within 1e-10 relative of the oracle's wherever
  |R_kk| >= 1e-5 |R_00| up to the divergence; backward error and
  orthogonality <= 50 n eps.  The oracle's run takes about half a minute on
  the host cores.
* The mixed-precision workload: 32768 x 8192, rank 512 (Gaussian factors,
  sigma_i = 1/i) + 1e-8 N(0,1), factored with the FP64, the FP32 and the
  BF16 sketch: backward error and orthogonality <= 50 n eps for each, and
  sigma_k(A) / |R_kk| at k = 512 within 2x of the FP64 sketch's, with
  sigma_k(A) from the exact SVD of the noise-free core (Weyl: the noise
  moves it by at most 1e-8 (sqrt(m) + sqrt(n)), i.e. ~0.14% here).
"""
import os

import numpy as np
import pytest

import oracle
from hqrrp_amd import matgen
from tests.helpers import check_pivots

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = 2.0 ** -52


def _factor(handle, At, m, n, sketch_order=False, **kw):
    import hqrrp_amd

    jp = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
    tau = torch.zeros(max(1, min(m, n)), dtype=torch.float64, device="cuda")
    sc = torch.zeros(max(1, min(m, n)), dtype=torch.int32, device="cuda")
    if sketch_order:
        kw["trace"] = {"sketch_cols": sc}
    hqrrp_amd.hqrrp_device(handle, At, jp, tau, m, n, **kw)
    handle.synchronize()
    torch.cuda.synchronize()
    return jp, tau, sc


def _backward(handle, A0t, At, jp, tau, m, n):
    """(||AP - QR||_F / ||A||_F, max|Q'Q - I|, R) on the device."""
    import hqrrp_amd

    k = min(m, n)
    Qt = hqrrp_amd.orgqr_device(handle, At, tau, m, k, k)
    handle.synchronize()
    torch.cuda.synchronize()
    R = torch.triu(At.t()[:k, :])
    AP = A0t[(jp.long() - 1)].t()
    res = (torch.linalg.matrix_norm(AP - Qt.t() @ R)
           / torch.linalg.matrix_norm(AP)).item()
    del AP
    orth = (Qt @ Qt.t() - torch.eye(k, dtype=torch.float64, device="cuda")
            ).abs().max().item()
    return res, orth, R


@pytest.mark.timeout(900)
def test_headline_pivots_match_oracle(handle):
    m = n = 16384
    nb, pp = 64, 10
    print(f"headline {m}x{n} graded: device factorisation", flush=True)
    A0t = matgen.graded_device(m, n, seed=0)
    At = A0t.clone()
    jp, tau, sc = _factor(handle, At, m, n, sketch_order=True, nb=nb, pp=pp,
                          panel_pivoting=1)
    res, orth, R = _backward(handle, A0t, At, jp, tau, m, n)
    print(f"residual {res:.2e} orthogonality {orth:.2e}", flush=True)
    assert res <= 50 * n * EPS and orth <= 50 * n * EPS
    jg = jp.cpu().numpy()
    sg = sc.cpu().numpy()
    dg = torch.diagonal(R).abs().cpu().numpy()
    A_host = np.asfortranarray(A0t.cpu().numpy().T)
    del At, R
    torch.cuda.empty_cache()
    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    print(f"oracle on {threads} host threads", flush=True)
    Ar, jr, _, ts, tp, sr = oracle.hqrrp(A_host, nb, pp, 1, trace=True)
    del A_host
    div = check_pivots(jg, jr, ts, tp, nb, sketch_gpu=sg, sketch_ref=sr,
                       label=f"headline {m}x{n}")
    dr = np.abs(np.diag(Ar))
    upto = n if div is None else div
    keep = dr[:upto] >= 1e-5 * dr[0]
    rel = np.abs(dg[:upto] - dr[:upto])[keep] / dr[:upto][keep]
    print(f"pivots equal up to {upto} of {n}; |diag R| checked on "
          f"{int(keep.sum())} entries, max relative difference "
          f"{rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-10


@pytest.mark.timeout(600)
def test_mixed_sketch_at_size(handle):
    m, n, k = 32768, 8192, 512
    A0t, s_core = matgen.lowrank_device(m, n, rank=k, seed=0, core_sigma=True)
    # sigma_k(A) within noise (sqrt(m) + sqrt(n)) of the core's (Weyl).
    sig_k = s_core[k - 1].item()
    weyl = 1e-8 * (m ** 0.5 + n ** 0.5) * 1.1
    assert weyl < 1e-2 * sig_k
    ratios = {}
    for prec, name in ((0, "fp64"), (1, "fp32"), (2, "bf16")):
        print(f"mixed workload {m}x{n} rank {k}: {name} sketch", flush=True)
        At = A0t.clone()
        jp, tau, _ = _factor(handle, At, m, n, nb=64, pp=10,
                             panel_pivoting=1, sketch_prec=prec)
        assert torch.equal(torch.sort(jp.long()).values,
                           torch.arange(1, n + 1, device="cuda"))
        res, orth, R = _backward(handle, A0t, At, jp, tau, m, n)
        rkk = R[k - 1, k - 1].abs().item()
        ratios[name] = sig_k / rkk
        print(f"{name}: residual {res:.2e} orthogonality {orth:.2e} "
              f"sigma_k/|R_kk| {ratios[name]:.3f} (|R_k+1,k+1| "
              f"{R[k, k].abs().item():.2e})", flush=True)
        assert res <= 50 * n * EPS and orth <= 50 * n * EPS
        del At, R
        torch.cuda.empty_cache()
    for name in ("fp32", "bf16"):
        assert 0.5 <= ratios[name] / ratios["fp64"] <= 2.0, ratios
