"""Batched persistent HQRRP of many small matrices, through the C ABI
(hqrrp_batch_create / hqrrp_batch_factor_dev).

Each matrix of a batch must be factored as the reference's
NoFLA_HQRRP_WY_blk_var4 factors it alone with the same nb and pp (the
reference has no batched entry; the oracle restates the single-matrix
driver):

* jpvt equal to the oracle's, or first diverging at a documented near-tie
  (relative gap of the two largest partial norms < 1e-12);
* ||AP - QR||/||A|| and ||Q'Q - I|| <= 50 max(m, n) eps;
* |diag R| within 1e-10 relative of the oracle's wherever |R_kk| >= 1e-5
  |R_00|, up to the first divergence.

Plus edge shapes (m < n, one row or column, nb > n, empty matrices),
oversampling 0 and 10, no panel pivoting, rank-deficient and tied columns, a
batch of the benchmark's distribution, graph capture (replays equal to the
eager run bit for bit) and run-to-run determinism.
"""
import numpy as np
import pytest

import oracle
from hqrrp_amd import matgen
from tests.helpers import check_pivots

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = np.finfo(np.float64).eps


def run_batch(handle, mats, nb, pp=8, pv=1):
    import hqrrp_amd

    m = [a.shape[0] for a in mats]
    n = [a.shape[1] for a in mats]
    pk = hqrrp_amd.PackedBatch(m, n)
    pk.load(mats)
    plan = hqrrp_amd.BatchPlan(handle, m, n, nb=nb, pp=pp, panel_pivoting=pv)
    plan.factor(handle, pk.A_ptrs, pk.jpvt_ptrs, pk.tau_ptrs)
    handle.synchronize()
    torch.cuda.synchronize()
    out = pk.unpack()
    nbs = plan.nb
    plan.close()
    return out, nbs


def check_one(A, got, nb, pp, pv, tag):
    Ag, jg, tg = got
    m, n = A.shape
    k = min(m, n)
    if k == 0:
        return
    Ar, jr, tr, ts, tp, sr = oracle.hqrrp(A, nb, pp, pv, trace=True)
    div = check_pivots(jg, jr, ts, tp, nb, sketch_ref=sr, panel_pivoting=pv,
                       label=tag)
    # Only the rank-deficient family (exactly zero and repeated columns:
    # candidates tie exactly once the rank is exhausted) may diverge at all.
    assert div is None or tag.startswith("rankdef"), (tag, div)
    res, orth = oracle.residuals(A, Ag, jg, tg)
    bound = 50 * max(m, n) * EPS
    assert res <= bound and orth <= bound, (tag, res, orth)
    upto = k if div is None else div
    if upto:
        dr = np.abs(np.diag(Ar)[:upto])
        dg = np.abs(np.diag(Ag)[:upto])
        keep = (dr > 0) & (dr >= 1e-5 * dr[0])
        if keep.any():
            rel = np.abs(dg - dr)[keep] / dr[keep]
            assert rel.max() <= 1e-10, (tag, rel.max())


def family(kind, m, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "graded":
        return matgen.graded(m, n, seed, decades=8.0)
    if kind == "gaussian":
        return np.asfortranarray(rng.standard_normal((m, n)))
    if kind == "rankdef":  # zero and repeated columns, rank ~ n/3
        k = max(1, min(m, n) // 3)
        A = rng.standard_normal((m, k)) @ rng.standard_normal((k, n))
        A[:, ::5] = 0.0
        if n > 3:
            A[:, 1] = A[:, 2]
        return np.asfortranarray(A)
    raise ValueError(kind)


SHAPES = [(64, 32), (100, 100), (300, 200), (512, 512), (37, 80), (1, 5),
          (5, 1), (129, 65), (511, 257), (200, 17), (16, 16), (8, 3),
          (256, 256), (73, 73), (450, 300), (64, 64), (2, 2), (512, 33)]


@pytest.mark.parametrize("nb", [16, 32, 64, 20])
def test_batched_parity(handle, nb):
    print(f"batched parity nb={nb}", flush=True)
    mats, tags = [], []
    for i, (m, n) in enumerate(SHAPES):
        kind = ("graded", "gaussian", "rankdef")[i % 3]
        mats.append(family(kind, m, n, 100 + i))
        tags.append(f"{kind} {m}x{n}")
    out, nbs = run_batch(handle, mats, nb)
    assert (nbs == nb).all()
    for A, got, tag in zip(mats, out, tags):
        check_one(A, got, nb, 8, 1, tag)


@pytest.mark.parametrize("pp,pv", [(0, 1), (10, 1), (8, 0)])
def test_batched_oversampling_and_panel(handle, pp, pv):
    print(f"batched pp={pp} pv={pv}", flush=True)
    shapes = [(200, 150), (512, 400), (90, 90), (40, 70)]
    mats = [family("graded", m, n, 7 + i) for i, (m, n) in enumerate(shapes)]
    out, _ = run_batch(handle, mats, 32, pp=pp, pv=pv)
    for A, got in zip(mats, out):
        check_one(A, got, 32, pp, pv, f"{A.shape} pp={pp} pv={pv}")


def test_batched_workload_distribution(handle):
    """256 matrices of the benchmark's distribution with the default block
    sizes: backward error of every matrix, oracle parity of a sample."""
    print("batched workload 256", flush=True)
    m, n = matgen.batch_shapes(256, seed=3)
    mats = matgen.batch_matrices(m, n, seed=3)
    out, nbs = run_batch(handle, mats, None)
    assert set(np.unique(nbs)) <= {16, 32, 64}
    for i, (A, got) in enumerate(zip(mats, out)):
        if i % 8 == 0:
            check_one(A, got, int(nbs[i]), 8, 1, f"workload {i}")
        else:
            res, orth = oracle.residuals(A, got[0], got[1], got[2])
            bound = 50 * max(A.shape) * EPS
            assert res <= bound and orth <= bound, (i, res, orth)


def test_batched_empty_and_degenerate(handle):
    import hqrrp_amd

    out, _ = run_batch(handle, [], 32)
    assert out == []
    mats = [np.zeros((0, 4)), np.zeros((3, 0)), np.zeros((6, 4)),
            np.ones((1, 1)) * 3.0]
    out, _ = run_batch(handle, mats, 16)
    assert out[2][1].tolist() == [1, 2, 3, 4]  # all-zero: no swaps, tau = 0
    assert np.all(out[2][2] == 0.0)
    assert out[3][0][0, 0] == 3.0 and out[3][2][0] == 0.0
    with pytest.raises(hqrrp_amd.HQRRPError):
        hqrrp_amd.BatchPlan(handle, [513], [10], nb=32)
    with pytest.raises(hqrrp_amd.HQRRPError):
        hqrrp_amd.BatchPlan(handle, [100], [10], nb=65)
    with pytest.raises(hqrrp_amd.HQRRPError):
        hqrrp_amd.BatchPlan(handle, [100], [10], nb=32, pp=11)


def test_batched_deterministic(handle):
    m, n = matgen.batch_shapes(64, seed=5)
    mats = matgen.batch_matrices(m, n, seed=5)
    a, _ = run_batch(handle, mats, 32)
    b, _ = run_batch(handle, mats, 32)
    for x, y in zip(a, b):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)


def test_batched_graph_capture(handle):
    """The factorisation captured in a graph: three replays (inputs reloaded
    eagerly in between) equal the eager run bit for bit."""
    import hqrrp_amd

    m, n = matgen.batch_shapes(48, seed=9)
    mats = matgen.batch_matrices(m, n, seed=9)
    pk = hqrrp_amd.PackedBatch(m, n)
    plan = hqrrp_amd.BatchPlan(handle, m, n, nb=32, pp=8)
    pk.load(mats)
    plan.factor(handle, pk.A_ptrs, pk.jpvt_ptrs, pk.tau_ptrs)
    handle.synchronize()
    eager = pk.unpack()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    pk.load(mats)
    pk.reset_pivots()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        handle.use_stream(s)
        plan.factor(handle, pk.A_ptrs, pk.jpvt_ptrs, pk.tau_ptrs)
    handle.use_stream(torch.cuda.current_stream())
    torch.cuda.synchronize()
    for rep in range(3):
        pk.load(mats)
        pk.reset_pivots()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = pk.unpack()
        for x, y in zip(eager, got):
            for u, v in zip(x, y):
                assert np.array_equal(u, v), rep
    plan.close()


def test_batched_autotune(handle):
    """The autotuner fills every bucket with a candidate block size, and a
    plan without explicit block sizes takes them from the table."""
    import hqrrp_amd

    table = handle.batch_tune(pp=8, panel_pivoting=1)
    assert set(np.unique(table)) <= {16, 32, 64}
    m = [64, 300, 512, 100]
    n = [32, 250, 512, 100]
    plan = hqrrp_amd.BatchPlan(handle, m, n, pp=8)
    want = [table[(mi - 1) // 64, (ni - 1) // 64] for mi, ni in zip(m, n)]
    assert plan.nb.tolist() == want
    plan.close()


@pytest.mark.timeout(600)
def test_batched_benchmark_launch_in_graph(handle):
    """The benchmark's whole 8192-matrix batch (seed 0, default block sizes,
    pp = 8), factored by a graph replay: two replays equal the eager launch
    bit for bit; every 64th matrix meets the backward-error bar and every
    512th the oracle's pivots."""
    import os

    import hqrrp_amd

    count = 8192
    print(f"batched benchmark launch {count}", flush=True)
    m, n = matgen.batch_shapes(count, seed=0)
    mats = matgen.batch_matrices(m, n, seed=0,
                                 workers=min(16, os.cpu_count() or 1))
    plan = hqrrp_amd.BatchPlan(handle, m, n, nb=None, pp=8)
    nbs = plan.nb
    src = hqrrp_amd.PackedBatch(m, n)
    src.load(mats)
    work = hqrrp_amd.PackedBatch(m, n)
    ident = src.jpvt.clone()

    def step():
        work.A.copy_(src.A)
        work.jpvt.copy_(ident)
        plan.factor(handle, work.A_ptrs, work.jpvt_ptrs, work.tau_ptrs)

    step()
    handle.synchronize()
    torch.cuda.synchronize()
    eagerA, eagerJ, eagerT = (work.A.clone(), work.jpvt.clone(),
                              work.tau.clone())
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        handle.use_stream(s)
        step()
    handle.use_stream(torch.cuda.current_stream())
    for rep in range(2):
        work.A.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(work.A, eagerA), rep
        assert torch.equal(work.jpvt, eagerJ), rep
        assert torch.equal(work.tau, eagerT), rep
    out = work.unpack()
    for i in range(0, count, 64):
        A = mats[i]
        if i % 512 == 0:
            check_one(A, out[i], int(nbs[i]), 8, 1, f"workload {i}")
        else:
            res, orth = oracle.residuals(A, *out[i])
            bound = 50 * max(A.shape) * EPS
            assert res <= bound and orth <= bound, (i, res, orth)
    plan.close()
