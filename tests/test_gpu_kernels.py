"""Parity of each HIP kernel with the oracle's restatement of the same
reference helper (NoFLA_HQRRP_WY_blk_var4.c), called through the C ABI.

Integer outputs (pivots) must match exactly; floating point within a few
hundred ulps of the operands' scale (the kernels sum in a different order
than the host BLAS)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from tests.helpers import moves_to_perm, swaps_to_perm

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    """numpy m x n -> device buffer holding it column-major."""
    import hqrrp_amd

    return hqrrp_amd.to_colmajor(a)


def host(t, m, n):
    return np.asfortranarray(t.detach().cpu().numpy().reshape(n, m).T)


def vp(t):
    return C.c_void_p(t.data_ptr())


def lib():
    import hqrrp_amd

    return hqrrp_amd.lib()


def sync(handle):
    handle.synchronize()
    torch.cuda.synchronize()


# ------------------------------------------------------------------ GEMM
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (74, 130, 1000), (5, 7, 3),
                                   (200, 33, 4097), (1, 1, 1)])
def test_dgemm_matches_numpy(handle, ta, tb, M, N, K):
    print(f"dgemm ta={ta} tb={tb} M={M} N={N} K={K}", flush=True)
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    Cm = rng.standard_normal((M, N))
    ref = 0.5 * (A.T if ta else A) @ (B.T if tb else B) - 2.0 * Cm
    dA, dB, dC = dev(A), dev(B), dev(Cm)
    rc = lib().hqrrp_dgemm_dev(handle.ptr, ta, tb, M, N, K, 0.5, vp(dA),
                               A.shape[0], vp(dB), B.shape[0], -2.0, vp(dC), M)
    assert rc == 0
    sync(handle)
    got = host(dC, M, N)
    scale = np.abs(A).max() * np.abs(B).max() * K + 2 * np.abs(Cm).max()
    assert np.max(np.abs(got - ref)) <= 1e-14 * scale


# ------------------------------------------------------ sketch pivoting
@pytest.mark.parametrize("r,ncols,stages,seed", [
    (74, 2000, 64, 0), (74, 300, 64, 1), (24, 1000, 16, 2), (40, 700, 32, 3),
    (74, 64, 64, 4), (74, 5000, 64, 5), (10, 257, 8, 6)])
def test_sketch_pivots_exact(handle, r, ncols, stages, seed):
    print(f"sketch r={r} ncols={ncols} stages={stages}", flush=True)
    rng = np.random.default_rng(seed)
    Y = rng.standard_normal((r, ncols)) * (1.0 + rng.random(ncols))
    ref = oracle.qrpmod(Y, pivoting=1, num_stages=stages, trace=True)
    dY = dev(Y)
    piv = torch.zeros(stages, dtype=torch.int32, device="cuda")
    moves = torch.full((1 + 4 * stages,), -7, dtype=torch.int32, device="cuda")
    rc = lib().hqrrp_sketch_qrcp_dev(handle.ptr, r, ncols, vp(dY), r, stages,
                                     vp(piv), vp(moves))
    assert rc == 0
    sync(handle)
    piv = piv.cpu().numpy()
    perm_swaps = swaps_to_perm(piv, ncols)
    perm_moves = moves_to_perm(moves.cpu().numpy(), ncols)
    assert np.array_equal(perm_swaps, perm_moves)
    assert np.array_equal(perm_swaps + 1, ref["p"])
    # Y is read only.
    assert np.array_equal(host(dY, r, ncols), Y)


def test_sketch_ties_take_first_position(handle):
    # Duplicate columns: equal norms at every step; idamax takes the first.
    rng = np.random.default_rng(11)
    base = rng.standard_normal((30, 40))
    Y = np.asfortranarray(np.concatenate([base, base, base[:, ::-1]], axis=1))
    ref = oracle.qrpmod(Y, pivoting=1, num_stages=20)
    dY = dev(Y)
    piv = torch.zeros(20, dtype=torch.int32, device="cuda")
    moves = torch.zeros(81, dtype=torch.int32, device="cuda")
    assert lib().hqrrp_sketch_qrcp_dev(handle.ptr, 30, 120, vp(dY), 30, 20,
                                       vp(piv), vp(moves)) == 0
    sync(handle)
    perm = swaps_to_perm(piv.cpu().numpy(), 120)
    # The first pivot is an exact tie between column c and its duplicate.
    assert perm[0] + 1 == ref["p"][0]


# ------------------------------------------------------------- panel QRCP
@pytest.mark.parametrize("mrows,b,pivoting,seed", [
    (2000, 64, 1, 0), (300, 64, 1, 1), (64, 64, 1, 2), (1000, 16, 1, 3),
    (777, 37, 0, 4), (5000, 64, 1, 5), (130, 5, 1, 6), (20000, 64, 1, 7)])
def test_panel_matches_oracle(handle, mrows, b, pivoting, seed):
    print(f"panel mrows={mrows} b={b} piv={pivoting}", flush=True)
    rng = np.random.default_rng(seed)
    P = rng.standard_normal((mrows, b)) * (1.0 + 3 * rng.random(b))
    ref = oracle.qrpmod(P, pivoting=pivoting, num_stages=-1, build_T=True)
    dP = dev(P)
    tau = torch.zeros(b, dtype=torch.float64, device="cuda")
    perm = torch.zeros(b, dtype=torch.int32, device="cuda")
    T = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    rc = lib().hqrrp_panel_qrcp_dev(handle.ptr, mrows, b, vp(dP), mrows,
                                    pivoting, vp(tau), vp(perm), vp(T), b)
    assert rc == 0
    sync(handle)
    assert np.array_equal(perm.cpu().numpy() + 1, ref["p"])
    got = host(dP, mrows, b)
    scale = np.abs(P).max() * np.sqrt(mrows)
    assert np.max(np.abs(got - ref["A"])) <= 1e-12 * scale
    assert np.max(np.abs(tau.cpu().numpy() - ref["tau"])) <= 1e-12
    Tg = T.cpu().numpy().reshape(b, b).T
    assert np.max(np.abs(Tg - ref["T"])) <= 1e-11 * max(1.0, np.abs(ref["T"]).max())


# ------------------------------------------------------------ dlarfb
@pytest.mark.parametrize("mrows,ncols,b", [(1000, 700, 64), (300, 5, 16),
                                           (64, 100, 64), (4000, 3000, 64),
                                           # V'C on the stream-K partition:
                                           # ragged groups, then a partial
                                           # last chunk with odd ld and b < 64
                                           (8192, 4100, 64), (8201, 4096, 40)])
def test_apply_qt_left(handle, mrows, ncols, b):
    print(f"apply_qt mrows={mrows} ncols={ncols} b={b}", flush=True)
    rng = np.random.default_rng(mrows + ncols)
    P = rng.standard_normal((mrows, b))
    fac = oracle.qrpmod(P, pivoting=1, num_stages=-1, build_T=True)
    Cm = rng.standard_normal((mrows, ncols))
    ref = oracle.apply_q_left_t(fac["A"], fac["T"], Cm)
    dV, dT, dC = dev(fac["A"]), dev(fac["T"]), dev(Cm)
    rc = lib().hqrrp_apply_qt_left_dev(handle.ptr, mrows, ncols, b, vp(dV),
                                       mrows, vp(dT), b, vp(dC), mrows)
    assert rc == 0
    sync(handle)
    got = host(dC, mrows, ncols)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(Cm).max() * np.sqrt(mrows)
    # V (with R above the diagonal) is untouched.
    assert np.array_equal(host(dV, mrows, b), fac["A"])


# ------------------------------------------------------- sketch downdate
@pytest.mark.parametrize("r,b,mU,nY2", [(74, 64, 1000, 900), (24, 16, 100, 50),
                                        (74, 64, 64 + 1, 1), (80, 64, 9000, 300),
                                        (40, 33, 777, 0), (90, 64, 500, 70),
                                        (74, 64, 16384, 129)])
def test_downdate_y(handle, r, b, mU, nY2):
    print(f"downdate r={r} b={b} mU={mU} nY2={nY2}", flush=True)
    rng = np.random.default_rng(r + mU)
    fac = oracle.qrpmod(rng.standard_normal((mU, b)), pivoting=1,
                        num_stages=-1, build_T=True)
    R12 = rng.standard_normal((b, nY2))
    Y2 = rng.standard_normal((r, nY2))
    G = rng.standard_normal((r, mU))
    Yr, Gr = oracle.downdate_y(fac["A"], R12, Y2, fac["T"], G)
    dU, dR, dT = dev(fac["A"]), dev(R12), dev(fac["T"])
    dY, dG = dev(Y2), dev(G)
    rc = lib().hqrrp_downdate_y_dev(handle.ptr, r, b, mU, nY2, vp(dU), mU,
                                    vp(dR), b, vp(dT), b, vp(dY), r, vp(dG), r)
    assert rc == 0
    sync(handle)
    s = np.sqrt(mU) * 10
    assert np.max(np.abs(host(dG, r, mU) - Gr)) <= 1e-12 * s
    if nY2:
        assert np.max(np.abs(host(dY, r, nY2) - Yr)) <= 1e-12 * s * np.sqrt(b)
