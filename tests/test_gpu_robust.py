"""Numerical range of the device paths.

* Power-of-two multiples of a matrix (2^+-600, 2^-700): the prescaling
  (launch_scale_pre) makes the factorisation exactly that of the unscaled
  matrix -- jpvt, reflectors and tau bitwise equal, R equal after the exact
  rescaling -- for HQRRP (FP64 and FP32 sketch), the classic QRCP and
  dgeqp4 with fixed columns.  Without it, squared norms of entries around
  1e-300 underflow to zero (and around 1e+300 overflow), and pivots would
  be noise.
* A non-power-of-two extreme scale (1e-300, 1e+300): pivots equal to the
  oracle's (LAPACK's scaled dnrm2 handles the range), backward error and
  orthogonality <= 50 n eps.
* NaN and Inf entries, including the exact bit pattern the persistent
  kernels use as their exchange sentinel: every entry point returns
  normally (no exchange time-out) with jpvt a permutation; a NaN column is
  pivoted after every finite one.
"""
import numpy as np
import pytest

import oracle
from hqrrp_amd import matgen
from tests.test_gpu_parity import run_gpu

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = np.finfo(np.float64).eps
SENTINEL = np.frombuffer(np.uint64(0x7FF4DEAD7FF4DEAD).tobytes(),
                         dtype=np.float64)[0]


def _split(Af):
    k = min(Af.shape)
    return np.triu(Af), np.tril(Af, -1)[:, :k]


@pytest.mark.parametrize("p", [600, -600, -900])
@pytest.mark.parametrize("prec", [0, 1])
def test_power_of_two_scaling_is_exact(handle, p, prec):
    A = matgen.graded(1200, 1000, 95, decades=6.0)
    s = 2.0 ** p
    a0, j0, t0 = run_gpu(handle, A, 64, 10, 1, sketch_prec=prec)
    a1, j1, t1 = run_gpu(handle, A * s, 64, 10, 1, sketch_prec=prec)
    assert np.array_equal(j0, j1)
    assert np.array_equal(t0, t1)
    R0, V0 = _split(a0)
    R1, V1 = _split(a1)
    assert np.array_equal(V0, V1)
    assert np.array_equal(R1 / s, R0)


def test_classic_and_dgeqp4_scaling_is_exact(handle):
    import hqrrp_amd

    s = 2.0 ** -700
    A = matgen.gaussian(300, 200, 96)

    def dev(M):
        At = hqrrp_amd.to_colmajor(M)
        jp = torch.arange(1, 201, dtype=torch.int32, device="cuda")
        tau = torch.zeros(200, dtype=torch.float64, device="cuda")
        hqrrp_amd.geqp3_device(handle, At, jp, tau, 300, 200)
        handle.synchronize()
        return (hqrrp_amd.from_colmajor(At), jp.cpu().numpy(),
                tau.cpu().numpy())

    a0, j0, t0 = dev(A)
    a1, j1, t1 = dev(A * s)
    assert np.array_equal(j0, j1) and np.array_equal(t0, t1)
    assert np.array_equal(np.tril(a0, -1), np.tril(a1, -1))
    assert np.array_equal(np.triu(a1) / s, np.triu(a0))
    # dgeqp4 with fixed columns (its HQRRP branch)
    B = matgen.gaussian(1200, 1000, 97)
    jin = np.zeros(1000, dtype=np.int32)
    jin[[3, 500, 999]] = 1
    b0, k0, u0, i0, _ = hqrrp_amd.dgeqp4(B, jpvt=jin)
    b1, k1, u1, i1, _ = hqrrp_amd.dgeqp4(B * s, jpvt=jin)
    assert i0 == 0 and i1 == 0
    assert np.array_equal(k0, k1) and np.array_equal(u0, u1)
    assert np.array_equal(np.tril(b0, -1), np.tril(b1, -1))
    assert np.array_equal(np.triu(b1) / s, np.triu(b0))


@pytest.mark.parametrize("scale", [1e-300, 1e300])
def test_extreme_scale_matches_oracle(handle, scale):
    A = matgen.gaussian(1100, 900, 98) * scale
    Ar, jr, tr = oracle.hqrrp(A, 64, 10, 1)
    Ag, jg, tg = run_gpu(handle, A, 64, 10, 1)
    assert np.array_equal(jg, jr)
    # The residual is formed on A / scale with R / scale (reflectors and tau
    # do not scale): numpy's norms of entries near 1e+-300 over/underflow.
    An = A / scale
    Agn = np.tril(Ag, -1) + np.triu(Ag) / scale
    res, orth = oracle.residuals(An, Agn, jg, tg)
    print(f"scale {scale:g}: residual {res:.2e} orthogonality {orth:.2e}",
          flush=True)
    assert res <= 50 * 1100 * EPS and orth <= 50 * 1100 * EPS


@pytest.mark.parametrize("kind", ["nan", "inf", "sentinel", "all_nan"])
def test_non_finite_inputs_return(handle, kind):
    """The factorisation of a matrix with NaN/Inf entries returns normally
    (LAPACK's dgeqp3 propagates them; the pivot order is then that of the
    rule above)."""
    import hqrrp_amd

    m, n = 900, 700
    A = matgen.gaussian(m, n, 99)
    if kind == "nan":
        A[10, 5] = np.nan
    elif kind == "inf":
        A[20, 9] = np.inf
    elif kind == "sentinel":
        A[0, :] = SENTINEL
        A[:, 3] = SENTINEL
    else:
        A[:] = np.nan
    print(f"non-finite input: {kind}", flush=True)
    for pv in (1, 0):
        Ag, jg, tg = run_gpu(handle, A, 64, 10, pv)
        assert sorted(jg.tolist()) == list(range(1, n + 1))
    if kind == "nan":
        # the NaN column is pivoted only after the finite candidates of the
        # first block's sketch and panel
        assert 6 not in jg[:64].tolist()
    # the other entry points with the same input
    Ag, jg, tg, info, _ = hqrrp_amd.dgeqp4(A)
    assert info == 0 and sorted(jg.tolist()) == list(range(1, n + 1))
    Ag, jg, tg, info, _ = hqrrp_amd.dgeqp4(A[:500, :400])
    assert info == 0 and sorted(jg.tolist()) == list(range(1, 401))
