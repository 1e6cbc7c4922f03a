"""The Cholesky-QR panel (k_cqr.hip) against the oracle's restatement of the
reference's panel QRCP (NoFLA_QRPmod_WY_unb_var4, NoFLA_HQRRP_WY_blk_var4.c:
774-886, with dlarft at 875-878), and its fallback to the per-step kernel.

Pivots must match exactly; values to rounding (the panel is factored by a
different sequence of operations: two Gram passes, a QRCP of the small R
factor and Householder reconstruction).  Panels the new path flags (nearly
dependent or zero columns, a column with nothing below its diagonal) are
factored by the per-step kernel, bitwise as with that kernel alone."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    import hqrrp_amd

    return hqrrp_amd.to_colmajor(a)


def host(t, m, n):
    return np.asfortranarray(t.detach().cpu().numpy().reshape(n, m).T)


def vp(t):
    return C.c_void_p(t.data_ptr())


def lib():
    import hqrrp_amd

    return hqrrp_amd.lib()


@pytest.fixture(scope="module")
def handles():
    """Two handles: the default panel (0) and the per-step kernel (1)."""
    import hqrrp_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    hs = []
    for alg in (0, 1):
        h = hqrrp_amd.Handle(0)
        h.use_stream(torch.cuda.current_stream())
        h.set_panel_alg(alg)
        hs.append(h)
    yield hs
    for h in hs:
        h.close()


def run_panel(h, P, pivoting):
    mrows, b = P.shape
    dP = dev(P)
    tau = torch.zeros(b, dtype=torch.float64, device="cuda")
    perm = torch.zeros(b, dtype=torch.int32, device="cuda")
    T = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    rc = lib().hqrrp_panel_qrcp_dev(h.ptr, mrows, b, vp(dP), mrows, pivoting,
                                    vp(tau), vp(perm), vp(T), b)
    assert rc == 0
    h.synchronize()
    torch.cuda.synchronize()
    return (host(dP, mrows, b), perm.cpu().numpy(), tau.cpu().numpy(),
            T.cpu().numpy().reshape(b, b).T)


def conditioned(rng, m, b, cond):
    U, _ = np.linalg.qr(rng.standard_normal((m, b)))
    V, _ = np.linalg.qr(rng.standard_normal((b, b)))
    return (U * np.logspace(0, -np.log10(cond), b)) @ V


CASES = [
    # (mrows, b, pivoting, kind)
    (2000, 64, 1, "gauss"), (300, 64, 1, "gauss"), (1000, 16, 1, "gauss"),
    (777, 37, 0, "gauss"), (5000, 64, 1, "gauss"), (130, 5, 1, "gauss"),
    (16384, 64, 1, "gauss"), (65536, 64, 1, "gauss"), (70, 64, 1, "gauss"),
    (4000, 64, 1, "cond1e4"), (4000, 64, 1, "graded"), (3001, 50, 0, "cond1e4"),
]


@pytest.mark.parametrize("mrows,b,pivoting,kind", CASES)
def test_cqr_panel_matches_oracle(handles, mrows, b, pivoting, kind):
    print(f"cqr panel mrows={mrows} b={b} piv={pivoting} {kind}", flush=True)
    rng = np.random.default_rng(mrows + 7 * b)
    if kind == "gauss":
        P = rng.standard_normal((mrows, b)) * (1.0 + 3 * rng.random(b))
    elif kind == "cond1e4":
        P = conditioned(rng, mrows, b, 1e4)
    else:  # columns of norms 1 .. 1e-6, nearly orthogonal
        P = rng.standard_normal((mrows, b)) * np.logspace(0, -6, b)
    ref = oracle.qrpmod(P, pivoting=pivoting, num_stages=-1, build_T=True)
    got, perm, tau, T = run_panel(handles[0], P, pivoting)
    assert np.array_equal(perm + 1, ref["p"])
    scale = np.abs(P).max() * np.sqrt(mrows)
    tol = 1e-12 if kind != "cond1e4" else 1e-10
    err = np.max(np.abs(got - ref["A"])) / scale
    print(f"  max |A - ref| / scale = {err:.2e}", flush=True)
    assert err <= tol
    assert np.max(np.abs(tau - ref["tau"])) <= 1e-12
    assert np.max(np.abs(T - ref["T"])) <= 100 * tol * max(1.0, np.abs(ref["T"]).max())
    # The new path ran (the per-step kernel's result differs in rounding).
    got1, perm1, _, _ = run_panel(handles[1], P, pivoting)
    assert np.array_equal(perm1, perm)
    assert not np.array_equal(got1, got)


FALLBACK = ["zero_col", "repeated_col", "square", "rank_deficient"]


@pytest.mark.parametrize("kind", FALLBACK)
def test_cqr_fallback_is_the_per_step_kernel(handles, kind):
    print(f"cqr fallback {kind}", flush=True)
    rng = np.random.default_rng(11)
    m, b = (64, 64) if kind == "square" else (1500, 64)
    P = rng.standard_normal((m, b))
    if kind == "zero_col":
        P[:, 7] = 0.0
    elif kind == "repeated_col":
        P[:, 20] = P[:, 3]
    elif kind == "rank_deficient":
        P[:, 32:] = P[:, :32] @ rng.standard_normal((32, 32))
    got0 = run_panel(handles[0], P, 1)
    got1 = run_panel(handles[1], P, 1)
    for a0, a1 in zip(got0, got1):
        assert np.array_equal(a0, a1)
    # Exact ties (a repeated column) and rank deficiency leave the pivot
    # order among equal or noise-level norms open.
    if kind in ("zero_col", "square"):
        ref = oracle.qrpmod(P, pivoting=1, num_stages=-1, build_T=True)
        assert np.array_equal(got0[1] + 1, ref["p"])


@pytest.mark.parametrize("m,n,la", [(3000, 3000, 1), (3000, 3000, 0),
                                    (5000, 1800, 1), (1500, 4000, 1)])
def test_cqr_driver_matches_per_step(handles, m, n, la):
    """Whole factorisations with either panel: the same pivots as the
    oracle's, backward error and orthogonality at working precision."""
    import hqrrp_amd

    print(f"cqr driver {m}x{n} lookahead={la}", flush=True)
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n))
    _, jr, _ = oracle.hqrrp(A, 64, 10, 1)
    eps = np.finfo(np.float64).eps
    for h in handles:
        lib().hqrrp_omega_carry_set(0, 0.0)
        dA = hqrrp_amd.to_colmajor(A, "cuda:0")
        jp = torch.arange(1, n + 1, dtype=torch.int32, device="cuda:0")
        tau = torch.zeros(min(m, n), dtype=torch.float64, device="cuda:0")
        hqrrp_amd.hqrrp_device(h, dA, jp, tau, m, n, nb=64, pp=10,
                               panel_pivoting=1, lookahead=la)
        h.synchronize()
        Ag = hqrrp_amd.from_colmajor(dA)
        jg = jp.cpu().numpy()
        assert np.array_equal(jg, jr)
        res, orth = oracle.residuals(A, Ag, jg, tau.cpu().numpy())
        print(f"  residual {res:.2e} orthogonality {orth:.2e}", flush=True)
        assert res < 50 * max(m, n) * eps and orth < 50 * max(m, n) * eps
