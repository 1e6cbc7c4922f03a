"""The drop-in parameter range beyond the persistent kernels: block sizes
nb > 64 (the reference's published b = 128, README.md:88), nb + pp > 80, more
than 65536 rows (one persistent panel launch) or columns (one persistent
sketch launch).  Those blocks run the generic block step (non-persistent
QRPmod steps for the sketch and the panel, T by blocks, the trailing update
and downdate as generic matrix-core products) -- still entirely on the GPU.
Each case against the oracle (NoFLA_HQRRP_WY_blk_var4.c:303-571 restated):
the pivot rule and value bars of tests/test_gpu_parity.py, backward error
and orthogonality <= 50 max(m,n) eps.
"""
import numpy as np
import pytest

import oracle
from hqrrp_amd import matgen
from tests.helpers import check_pivots
from tests.test_gpu_parity import run_gpu, sketched_columns

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = np.finfo(np.float64).eps


def _family(kind, m, n, seed):
    if kind == "gaussian":
        return matgen.gaussian(m, n, seed)
    return matgen.graded(m, n, seed, decades=6.0)


CASES = [
    # (m, n, nb, pp, panel_pivoting, family, seed)
    (1500, 1200, 128, 10, 1, "gaussian", 81),   # the published b = 128
    (1300, 1300, 128, 16, 1, "graded", 82),
    (1000, 900, 96, 16, 1, "gaussian", 83),    # nb + pp = 112
    (900, 800, 100, 5, 1, "graded", 84),       # nb not a power of two
    (1000, 700, 72, 16, 0, "gaussian", 85),    # no panel pivoting
    (70000, 1500, 64, 10, 1, "gaussian", 86),  # panels of > 65536 rows
    (1500, 70000, 64, 10, 1, "gaussian", 87),  # sketch of > 65536 columns
]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("m,n,nb,pp,pv,family,seed", CASES)
def test_generic_block_step_parity(handle, m, n, nb, pp, pv, family, seed):
    print(f"generic {family} {m}x{n} nb={nb} pp={pp} pv={pv}", flush=True)
    A = _family(family, m, n, seed)
    Ar, jr, tr, ts, tp, sr = oracle.hqrrp(A, nb, pp, pv, trace=True)
    Ag, jg, tg, sg = run_gpu(handle, A, nb, pp, pv, sketch_order=True)
    div = check_pivots(jg, jr, ts, tp, nb, sketch_gpu=sg, sketch_ref=sr,
                       panel_pivoting=pv, label=f"{family} {m}x{n} nb={nb}")
    assert div is None, f"pivots diverge at {div}"
    nsk = sketched_columns(m, n, nb)
    assert np.array_equal(sg[:nsk], sr[:nsk])
    res, orth = oracle.residuals(A, Ag, jg, tg)
    print(f"residual {res:.2e} orthogonality {orth:.2e}", flush=True)
    assert res <= 50 * max(m, n) * EPS and orth <= 50 * max(m, n) * EPS
    dr, dg = np.abs(np.diag(Ar)), np.abs(np.diag(Ag))
    keep = dr >= 1e-5 * dr[0]
    assert (np.abs(dg - dr)[keep] / dr[keep]).max() <= 1e-10


@pytest.mark.parametrize("prec", [1, 2])
def test_generic_mixed_sketch(handle, prec):
    """nb = 128 with an FP32 or BF16 sketch (generic sketch on the widened
    copy, generic FP32 downdate): still a backward-stable factorisation."""
    m, n = 1400, 1100
    A = matgen.graded(m, n, 88, decades=6.0)
    Ag, jg, tg = run_gpu(handle, A, 128, 10, 1, sketch_prec=prec)
    assert sorted(jg.tolist()) == list(range(1, n + 1))
    res, orth = oracle.residuals(A, Ag, jg, tg)
    assert res <= 50 * m * EPS and orth <= 50 * m * EPS


def test_generic_panel_component(handle):
    """hqrrp_panel_qrcp_dev with b = 128 (generic panel + blocked T): pivots
    equal to the oracle's QRPmod, values to rounding, T against dlarft."""
    import hqrrp_amd

    m, b = 3000, 128
    P = matgen.graded(m, b, 89, decades=4.0)
    ref = oracle.qrpmod(P, pivoting=1, num_stages=-1, build_T=True)
    dP = hqrrp_amd.to_colmajor(P)
    tau = torch.zeros(b, dtype=torch.float64, device="cuda")
    perm = torch.zeros(b, dtype=torch.int32, device="cuda")
    T = torch.zeros((b, b), dtype=torch.float64, device="cuda")
    hqrrp_amd.lib().hqrrp_panel_qrcp_dev(
        handle.ptr, m, b, hqrrp_amd.api._vp(dP), m, 1, hqrrp_amd.api._vp(tau),
        hqrrp_amd.api._vp(perm), hqrrp_amd.api._vp(T), b)
    handle.synchronize()
    torch.cuda.synchronize()
    assert np.array_equal(perm.cpu().numpy() + 1, ref["p"])
    Pg = hqrrp_amd.from_colmajor(dP)
    assert np.abs(Pg - ref["A"]).max() <= 1e-12 * max(1.0, np.abs(P).max())
    assert np.abs(tau.cpu().numpy() - ref["tau"]).max() <= 1e-12
    Tg = T.cpu().numpy().T  # (b, b) row-major buffer = column-major T
    assert np.abs(Tg - ref["T"]).max() <= 1e-11


def test_generic_sketch_component(handle):
    """hqrrp_sketch_qrcp_dev with r = 144 > 80 and 128 stages (generic
    sketch): the oracle QRPmod's pivots."""
    import hqrrp_amd
    from tests.helpers import moves_to_perm, swaps_to_perm

    r, ncols, stages = 144, 3000, 128
    Y = np.random.default_rng(90).standard_normal((r, ncols))
    ref = oracle.qrpmod(Y, pivoting=1, num_stages=stages)
    dY = hqrrp_amd.to_colmajor(Y)
    piv = torch.zeros(stages, dtype=torch.int32, device="cuda")
    mv = torch.zeros(1 + 4 * stages, dtype=torch.int32, device="cuda")
    rc = hqrrp_amd.lib().hqrrp_sketch_qrcp_dev(
        handle.ptr, r, ncols, hqrrp_amd.api._vp(dY), r, stages,
        hqrrp_amd.api._vp(piv), hqrrp_amd.api._vp(mv))
    assert rc == 0
    handle.synchronize()
    torch.cuda.synchronize()
    p = swaps_to_perm(piv.cpu().numpy(), ncols)
    assert np.array_equal(p[:stages] + 1, ref["p"][:stages])
    assert np.array_equal(moves_to_perm(mv.cpu().numpy(), ncols), p)
    # the sketch itself is read, not modified
    assert np.array_equal(hqrrp_amd.from_colmajor(dY), Y)


def test_dgeqp4_tall_fixed_columns(handle):
    """dgeqp4 on 70000 x 300 with three fixed columns (its dgeqp3 branch,
    min(m,n) < 1000): the fixed columns' QR runs on the generic panel."""
    import hqrrp_amd

    m, n = 70000, 300
    A = matgen.gaussian(m, n, 91)
    jin = np.zeros(n, dtype=np.int32)
    jin[[5, 17, 200]] = 1
    Ag, jg, tg, info, _ = hqrrp_amd.dgeqp4(A, jpvt=jin)
    Ar, jr, tr, info_r, _ = oracle.dgeqp4(A, jpvt=jin)
    assert info == 0 and info_r == 0
    assert sorted(jg[:3].tolist()) == [6, 18, 201]
    assert np.array_equal(jg, jr)
    res, orth = oracle.residuals(A, Ag, jg, tg)
    assert res <= 50 * m * EPS and orth <= 50 * m * EPS
