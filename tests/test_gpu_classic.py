"""Classic greedy QRCP on the device (hqrrp_geqp3_dev) and dgeqp4's small
branches, which the reference sends to LAPACK (min(m,n) < 250 -> dgeqpf,
< 1000 -> dgeqp3; NoFLA_HQRRP_WY_blk_var4.c:189-199).  The oracle is that
LAPACK (OpenBLAS dgeqp3 / dgeqpf; its dgeqp4 is bitwise equal to the compiled
reference front end):

* jpvt equal, or -- only for the cases listed in TIES -- first diverging at
  step i at a near-tie: the two candidate columns' partial norms at step i,
  read off the oracle's R (||R(i:k+1, k)|| is the partial norm after i steps
  of the column it placed at k), differ by less than 1e-12 |R_00|; every
  divergence is printed;
* ||AP - QR||/||A||, ||Q'Q - I|| <= 50 max(m,n) eps;
* |diag R| within 1e-10 |R_00| of the oracle's up to the divergence;
* the reference example driver's 7x5 case (fixed columns, dgeqpf branch)
  gives its recorded output.
"""
import os

import numpy as np
import pytest

import oracle
from hqrrp_amd import matgen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = np.finfo(np.float64).eps
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _matrix(family, m, n, seed):
    if family == "gaussian":
        return matgen.gaussian(m, n, seed)
    if family == "graded":
        return matgen.graded(m, n, seed)
    if family == "lowrank":
        return matgen.lowrank_noise(m, n, rank=min(m, n) // 3, seed=seed)
    A = matgen.gaussian(m, n, seed)          # "defective": zero and repeated
    A[:, 3] = 0.0
    A[:, 10] = A[:, 4]
    A[:, n - 1] = 2.0 * A[:, 7]
    return A


def _compare(A, Ag, jg, tg, Ar, jr, tr):
    m, n = A.shape
    k = min(m, n)
    jg, jr = np.asarray(jg), np.asarray(jr)
    scale = abs(Ar[0, 0])
    d = np.nonzero(jg != jr)[0]
    upto = k
    if d.size:
        i = int(d[0])
        kpos = int(np.nonzero(jr == jg[i])[0][0])
        assert kpos > i
        top = min(kpos + 1, m)
        gap = abs(Ar[i, i]) - np.linalg.norm(Ar[i:top, kpos])
        print(f"{m}x{n}: pivots diverge at {i}, oracle gap "
              f"{gap / scale:.3e} of |R_00|", flush=True)
        assert gap < 1e-12 * scale, (
            f"pivots diverge at {i}: oracle gap {gap / scale:.3e} of |R_00|")
        upto = i
    res, orth = oracle.residuals(A, Ag, jg, tg)
    assert res <= 50 * max(m, n) * EPS and orth <= 50 * max(m, n) * EPS
    dg, dr = np.abs(np.diag(Ag)[:upto]), np.abs(np.diag(Ar)[:upto])
    if upto:
        assert np.abs(dg - dr).max() <= 1e-10 * scale
    return None if not d.size else int(d[0])


def _run_dev(handle, A):
    import hqrrp_amd

    m, n = A.shape
    dA = hqrrp_amd.to_colmajor(A)
    jp = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
    tau = torch.zeros(max(1, min(m, n)), dtype=torch.float64, device="cuda")
    hqrrp_amd.geqp3_device(handle, dA, jp, tau, m, n)
    handle.synchronize()
    torch.cuda.synchronize()
    return (hqrrp_amd.from_colmajor(dA), jp.cpu().numpy(),
            tau.cpu().numpy()[:min(m, n)])


CASES = [
    (300, 200, "gaussian", 1), (200, 300, "gaussian", 2),
    (999, 999, "graded", 3), (1500, 120, "graded", 4),
    (64, 2000, "gaussian", 5), (257, 257, "lowrank", 6),
    (400, 300, "defective", 7), (1, 40, "gaussian", 8),
    (40, 1, "gaussian", 9), (2, 2, "gaussian", 10),
    # the one-launch path's limits (1024 rows, 2048 columns) and past them
    (1024, 700, "graded", 11), (700, 2048, "gaussian", 12),
    (100, 3000, "gaussian", 13), (1025, 300, "graded", 14),
]


# Cases whose pivots diverge from LAPACK's at a near-tie (the defective
# matrix's last columns are rounding noise: gap 6e-17 of |R_00| at step
# 297); every other case must match exactly.
TIES = {(400, 300, "defective")}


@pytest.mark.parametrize("m,n,family,seed", CASES)
def test_geqp3_device_matches_lapack(handle, m, n, family, seed):
    print(f"geqp3 {family} {m}x{n}", flush=True)
    A = _matrix(family, m, n, seed)
    Ag, jg, tg = _run_dev(handle, A)
    Ar, jr, tr, info = oracle.dgeqp3(A)
    assert info == 0
    div = _compare(A, Ag, jg, tg, Ar, jr, tr)
    if (m, n, family) not in TIES:
        assert div is None, f"pivots diverge at {div}"


@pytest.mark.parametrize("m,n,fixed", [
    (200, 150, [0, 5, 77]),          # dgeqpf branch
    (230, 600, [599, 1, 300]),       # dgeqpf branch, wide
    (600, 500, []),                  # dgeqp3 branch
    (700, 800, list(range(40, 90))),  # dgeqp3 branch, 50 fixed
    (50, 120, list(range(0, 120, 3)) + [1]),  # 41 fixed < m
    (30, 120, list(range(0, 80))),   # more fixed columns than rows
])
def test_dgeqp4_small_branches(handle, m, n, fixed):
    import hqrrp_amd

    print(f"dgeqp4 small {m}x{n} fixed={len(fixed)}", flush=True)
    A = matgen.gaussian(m, n, 31 + m)
    jin = np.zeros(n, dtype=np.int32)
    jin[fixed] = 1
    Ag, jg, tg, info, w0 = hqrrp_amd.dgeqp4(A, jpvt=jin)
    Ar, jr, tr, info_r, _ = oracle.dgeqp4(A, jpvt=jin)
    assert info == 0 and info_r == 0 and w0 == 3 * n + 1
    nf = min(len(set(fixed)), m)
    assert sorted(jg[:nf].tolist()) == sorted(jr[:nf].tolist())
    assert _compare(A, Ag, jg, tg, Ar, jr, tr) is None


def test_dgeqp4_small_matches_device_entry(handle):
    """No fixed columns: the host dgeqp4 is the device entry, bit for bit."""
    import hqrrp_amd

    A = matgen.graded(600, 500, 41)
    Ah, jh, th, info, _ = hqrrp_amd.dgeqp4(A)
    Ad, jd, td = _run_dev(handle, A)
    assert info == 0
    assert np.array_equal(jh, jd)
    assert np.array_equal(th, td)
    assert np.array_equal(Ah, Ad)


def test_simple_test_known_answer_on_gpu(handle):
    """The reference example driver's 7x5 integer matrix with fixed columns
    {2, 3, 5}: its printed pivots, tau and diag R, and the recorded outputs."""
    import hqrrp_amd

    d = np.load(os.path.join(GOLD, "simple_test_7x5.npz"))
    Ag, jg, tg, info, _ = hqrrp_amd.dgeqp4(d["A"], jpvt=d["jpvt_in"])
    assert info == 0
    assert list(jg) == [2, 3, 5, 1, 4]
    np.testing.assert_allclose(tg, [1.473286, 1.981109, 1.731670, 1.173021,
                                    1.384567], atol=5e-7)
    np.testing.assert_allclose(np.diag(Ag), [-29.58040, -8.851215, -11.74025,
                                             -4.940191, -4.915331], atol=5e-6)
    assert np.array_equal(jg, d["jpvt"])
    assert np.abs(tg - d["tau"]).max() <= 1e-13
    assert np.abs(Ag - d["A_out"]).max() <= 1e-12 * np.abs(d["A_out"]).max()
