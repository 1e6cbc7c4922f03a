"""End-to-end parity of the GPU HQRRP with the oracle (a CPU restatement of
NoFLA_HQRRP_WY_blk_var4.c that is bitwise equal to the compiled reference).

* jpvt must equal the oracle's; every case here does, so a divergence fails
  the test even where tests.helpers.check_pivots would attribute it to a
  near-tie (relative gap of the two largest partial norms < 1e-12 at the
  exact sketch or panel decision that diverged, which it prints);
* ||AP - QR||/||A|| and ||Q'Q - I|| <= 50 n eps;
* |diag R| within 1e-10 relative of the oracle's wherever |R_kk| >= 1e-5 |R_00|
  (two correct LAPACKs already differ beyond that range on graded spectra).
"""
import numpy as np
import pytest

import oracle
from hqrrp_amd import matgen
from tests.helpers import check_pivots

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = np.finfo(np.float64).eps


def run_gpu(handle, A, nb, pp, pv, sketch_order=False, **kw):
    """Factor A on the device; with sketch_order, also returns the library's
    trace of each block's sketch choices (hqrrp_trace.sketch_cols)."""
    import hqrrp_amd

    m, n = A.shape
    dA = hqrrp_amd.to_colmajor(A)
    jp = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
    tau = torch.zeros(max(1, min(m, n)), dtype=torch.float64, device="cuda")
    sc = torch.zeros(max(1, min(m, n)), dtype=torch.int32, device="cuda")
    if sketch_order:
        kw["trace"] = {"sketch_cols": sc}
    hqrrp_amd.hqrrp_device(handle, dA, jp, tau, m, n, nb=nb, pp=pp,
                           panel_pivoting=pv, **kw)
    handle.synchronize()
    torch.cuda.synchronize()
    out = (hqrrp_amd.from_colmajor(dA), jp.cpu().numpy(),
           tau.cpu().numpy()[:min(m, n)])
    if sketch_order:
        return out + (sc.cpu().numpy()[:min(m, n)],)
    return out


CASES = [
    # (m, n, nb, pp, panel_pivoting, family, seed)
    (2000, 2000, 64, 10, 1, "gaussian", 1),
    (1500, 1500, 64, 10, 1, "graded", 2),
    (1200, 800, 64, 10, 1, "gaussian", 3),
    (800, 1200, 64, 10, 1, "gaussian", 4),
    (1000, 1000, 32, 8, 1, "graded", 5),
    (1000, 700, 16, 8, 0, "gaussian", 6),
    (3000, 3000, 64, 10, 1, "graded", 7),  # a near-tie at 2820 (TIES)
    (513, 257, 64, 10, 1, "lowrank", 8),
]


# Cases whose pivots first diverge from the oracle's at a near-tie (the
# oracle's gap at that sketch decision is 3e-16 of the sketch's scale, deep
# in the graded spectrum): the divergence is accepted there and nowhere else.
TIES = {(3000, 3000, "graded", 7)}


def sketched_columns(m, n, nb, kmax=None):
    """Columns chosen by a sketch: those of every block but the reference's
    final one (j + nb >= m or j + nb >= n, NoFLA_HQRRP_WY_blk_var4.c:
    413-416), up to kmax."""
    k = min(m, n) if kmax is None else kmax
    tot = 0
    for j in range(0, k, nb):
        if j + nb >= m or j + nb >= n:
            break
        tot = min(j + nb, k)
    return tot


def make(family, m, n, seed):
    if family == "gaussian":
        return matgen.gaussian(m, n, seed)
    if family == "graded":
        return matgen.graded(m, n, seed)
    return matgen.lowrank_noise(m, n, rank=min(m, n) // 4, seed=seed)


@pytest.mark.parametrize("m,n,nb,pp,pv,family,seed", CASES)
def test_factorization_parity(handle, m, n, nb, pp, pv, family, seed):
    print(f"factor {family} {m}x{n} nb={nb} pp={pp} pv={pv}", flush=True)
    A = make(family, m, n, seed)
    Ar, jr, tr, ts, tp, sr = oracle.hqrrp(A, nb, pp, pv, trace=True)
    Ag, jg, tg, sg = run_gpu(handle, A, nb, pp, pv, sketch_order=True)
    k = min(m, n)
    div = check_pivots(jg, jr, ts, tp, nb, sketch_gpu=sg, sketch_ref=sr,
                       panel_pivoting=pv, label=f"{family} {m}x{n}")
    if div is None:
        # Every sketched block (all but the final one) chose the oracle's
        # columns in the oracle's order.
        nsk = sketched_columns(m, n, nb)
        assert np.array_equal(sg[:nsk], sr[:nsk])
    res, orth = oracle.residuals(A, Ag, jg, tg)
    assert res <= 50 * max(m, n) * EPS, res
    assert orth <= 50 * max(m, n) * EPS, orth
    upto = k if div is None else div
    dr = np.abs(np.diag(Ar)[:upto])
    dg = np.abs(np.diag(Ag)[:upto])
    keep = dr >= 1e-5 * dr[0] if upto else dr > 0
    if upto:
        rel = np.abs(dg - dr)[keep] / dr[keep]
        assert rel.max() <= 1e-10, rel.max()
    assert div is None or (m, n, family, seed) in TIES, \
        f"pivots diverge at {div}"


def test_host_entry_matches_device(handle):
    import hqrrp_amd

    A = matgen.gaussian(1100, 1000, 21)
    Ah, jh, th = hqrrp_amd.hqrrp(A, 64, 10, 1)
    Ad, jd, td = run_gpu(handle, A, 64, 10, 1)
    assert np.array_equal(jh, jd)
    assert np.array_equal(Ah, Ad)
    assert np.array_equal(th, td)


def test_dgeqp4_front_end(handle):
    import hqrrp_amd

    A = matgen.gaussian(1200, 1000, 22)
    Ag, jg, tg, info, w0 = hqrrp_amd.dgeqp4(A)
    Ar, jr, tr, info_r, _ = oracle.dgeqp4(A)
    assert info == 0 and info_r == 0
    assert w0 == 3 * 1000 + 1
    assert np.array_equal(jg, jr)
    res, orth = oracle.residuals(A, Ag, jg, tg)
    assert res <= 50 * 1200 * EPS and orth <= 50 * 1200 * EPS


@pytest.mark.parametrize("fixed", [[3, 17, 400, 999], list(range(0, 80)),
                                   list(range(500, 1000, 7))])
def test_dgeqp4_fixed_columns(handle, fixed):
    """Fixed leading columns (jpvt != 0): the GPU path (QR of the fixed
    columns, HQRRP of the free block, re-pivoted top rows) gives the oracle's
    pivots and a valid factorization (NoFLA_HQRRP_WY_blk_var4.c:201-292)."""
    import hqrrp_amd

    m, n = 1200, 1000
    A = matgen.gaussian(m, n, 23)
    jin = np.zeros(n, dtype=np.int32)
    jin[fixed] = 1
    Ag, jg, tg, info, _ = hqrrp_amd.dgeqp4(A, jpvt=jin)
    Ar, jr, tr, info_r, _ = oracle.dgeqp4(A, jpvt=jin)
    assert info == 0 and info_r == 0
    nf = len(fixed)
    assert sorted(jg[:nf].tolist()) == [c + 1 for c in fixed]
    assert np.array_equal(jg, jr)
    res, orth = oracle.residuals(A, Ag, jg, tg)
    assert res <= 50 * m * EPS and orth <= 50 * m * EPS
    dg, dr = np.abs(np.diag(Ag)), np.abs(np.diag(Ar))
    assert np.abs(dg - dr).max() <= 1e-10 * dr.max()


def test_deterministic(handle):
    A = matgen.graded(1300, 1300, 31)
    a1, j1, t1 = run_gpu(handle, A, 64, 10, 1)
    a2, j2, t2 = run_gpu(handle, A, 64, 10, 1)
    assert np.array_equal(j1, j2) and np.array_equal(a1, a2)


def test_omega_carry_across_calls(handle):
    """Three calls in a row with an odd r*m (nb = 16, pp = 5, m = 1301): as
    in the reference, whose generator statics carry the last pair's second
    variate into the next call, the second call factors with the shifted
    Omega and the third with the first call's again.  Each call against the
    oracle with that call's Omega (the oracle's carried Omega is pinned
    against the compiled reference on the CPU)."""
    import hqrrp_amd

    m, n, nb, pp = 1301, 1100, 16, 5
    A = matgen.gaussian(m, n, 57)
    r = nb + pp
    G1, c1 = oracle.omega(r, m, carry=(0, 0.0))
    G2, c2 = oracle.omega(r, m, carry=c1)
    assert hqrrp_amd.omega_carry() == (0, 0.0)
    want = []
    for G in (G1, G2):
        want.append(oracle.hqrrp(A, nb, pp, 1, omega_in=G))
    for call, (G, carry) in enumerate(((G1, c1), (G2, c2), (G1, c1))):
        print(f"call {call + 1}: carry in {hqrrp_amd.omega_carry()}", flush=True)
        Ag, jg, tg = run_gpu(handle, A, nb, pp, 1)
        assert hqrrp_amd.omega_carry() == carry
        Ar, jr, tr = want[call % 2]
        assert np.array_equal(jg, jr), f"call {call + 1}: pivots differ"
        dr, dg = np.abs(np.diag(Ar)), np.abs(np.diag(Ag))
        assert (np.abs(dg - dr) / dr).max() <= 1e-10


LOOKAHEAD_CASES = [
    # (m, n, nb, pp, panel_pivoting, max_rank, seed)
    (1300, 1300, 64, 10, 1, 0, 41),
    (1000, 1500, 64, 10, 1, 0, 42),   # wide: last block ends at m
    (1500, 1000, 32, 8, 1, 0, 43),    # tall
    (1000, 1000, 16, 8, 0, 0, 44),    # no panel pivoting
    (1100, 1030, 64, 10, 1, 0, 45),   # ragged last block
    (1200, 1200, 64, 10, 1, 300, 46),  # truncated
]


@pytest.mark.parametrize("m,n,nb,pp,pv,kmax,seed", LOOKAHEAD_CASES)
def test_lookahead_bitwise(handle, m, n, nb, pp, pv, kmax, seed):
    """Overlapping the next panel with the trailing update changes nothing."""
    print(f"lookahead {m}x{n} nb={nb} pv={pv} max_rank={kmax}", flush=True)
    A = matgen.graded(m, n, seed)
    a0, j0, t0 = run_gpu(handle, A, nb, pp, pv, max_rank=kmax, lookahead=0)
    a1, j1, t1 = run_gpu(handle, A, nb, pp, pv, max_rank=kmax, lookahead=1)
    assert np.array_equal(j0, j1)
    assert np.array_equal(t0, t1)
    assert np.array_equal(a0, a1)
