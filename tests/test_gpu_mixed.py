"""Mixed-precision sketch (FP32 sketch, and bf16 operands in the sketch GEMM
with FP32 accumulation): no reference counterpart, pinned by the properties
it must keep.

* The factorization itself stays FP64 Householder, so the backward error
  ||AP - QR||/||A|| and ||Q'Q - I|| meet the same 50 n eps bar as the FP64
  path.
* Rank revelation: on a rank-k matrix plus 1e-8 noise (the mixed-precision
  workload's family), sigma_k / |R_kk| at the numerical rank is within 2x of
  the FP64 sketch's ratio, and |R_kk| collapses after k.
* Look-ahead on and off give bitwise the same result.
"""
import numpy as np
import pytest

import oracle
from tests.test_gpu_parity import run_gpu

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = np.finfo(np.float64).eps
PREC = {"fp64": 0, "fp32": 1, "bf16": 2}


def lowrank(m, n, k, seed, noise=1e-8):
    rng = np.random.default_rng(seed)
    L = rng.standard_normal((m, k))
    R = rng.standard_normal((n, k)) / np.arange(1, k + 1)
    return L @ R.T + noise * rng.standard_normal((m, n))


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("m,n,k", [(2000, 1500, 200), (3000, 1000, 300)])
def test_mixed_backward_error_and_rank(handle, prec, m, n, k):
    print(f"mixed {prec} {m}x{n} rank {k}", flush=True)
    A = lowrank(m, n, k, 11)
    sv = np.linalg.svd(A, compute_uv=False)
    out = {}
    for p in ("fp64", prec):
        Ag, jg, tg = run_gpu(handle, A, 64, 10, 1, sketch_prec=PREC[p])
        res, orth = oracle.residuals(A, Ag, jg, tg)
        assert res <= 50 * n * EPS and orth <= 50 * n * EPS, (p, res, orth)
        out[p] = np.abs(np.diag(Ag))
    r64 = sv[k - 1] / out["fp64"][k - 1]
    rmx = sv[k - 1] / out[prec][k - 1]
    assert rmx <= 2.0 * r64 and rmx >= 0.5 * r64, (r64, rmx)
    # the numerical rank shows: |R_kk| drops by orders of magnitude after k
    assert out[prec][k] <= 1e-4 * out[prec][k - 1]


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_mixed_graded_backward_error(handle, prec):
    from hqrrp_amd import matgen

    A = matgen.graded(1500, 1500, 12)
    Ag, jg, tg = run_gpu(handle, A, 64, 10, 1, sketch_prec=PREC[prec])
    res, orth = oracle.residuals(A, Ag, jg, tg)
    assert res <= 50 * 1500 * EPS and orth <= 50 * 1500 * EPS
    assert sorted(jg.tolist()) == list(range(1, 1501))


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_mixed_lookahead_bitwise(handle, prec):
    A = lowrank(1300, 1100, 150, 13)
    a0, j0, t0 = run_gpu(handle, A, 64, 10, 1, sketch_prec=PREC[prec],
                         lookahead=0)
    a1, j1, t1 = run_gpu(handle, A, 64, 10, 1, sketch_prec=PREC[prec],
                         lookahead=1)
    assert np.array_equal(j0, j1)
    assert np.array_equal(a0, a1) and np.array_equal(t0, t1)
