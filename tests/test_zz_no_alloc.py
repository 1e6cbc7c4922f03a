"""hqrrp_geqp3_noalloc_dev, the allocation-free variant of hqrrp_geqp3_dev:
called through the C ABI, its results equal hqrrp_geqp3_dev's bit for bit
on two calls in a row (the second smaller, so stale workspace from the
first would show), on both kernel paths (one launch per step for m <= 1024,
n <= 2048; two launches per step beyond); a shape whose workspace exceeds
the static scratch is refused with HQRRP_ERR_UNSUPPORTED and leaves its
arguments untouched."""
import ctypes as C

import numpy as np
import pytest

from hqrrp_amd import matgen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _vp(t):
    return C.c_void_p(t.data_ptr())


def _call(handle, name, A):
    import hqrrp_amd

    m, n = A.shape
    At = hqrrp_amd.to_colmajor(A)
    jp = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
    tau = torch.zeros(max(1, min(m, n)), dtype=torch.float64, device="cuda")
    rc = getattr(hqrrp_amd.lib(), name)(handle.ptr, m, n, _vp(At), m, _vp(jp),
                                        _vp(tau))
    handle.synchronize()
    torch.cuda.synchronize()
    return rc, At.cpu(), jp.cpu(), tau.cpu()


@pytest.mark.parametrize("shapes", [((700, 600), (300, 200)),
                                    ((1500, 300), (1100, 120))])
def test_noalloc_equals_geqp3_dev(handle, shapes):
    for i, (m, n) in enumerate(shapes):
        print(f"noalloc call {i}: {m}x{n}", flush=True)
        A = matgen.graded(m, n, seed=60 + i, decades=8.0)
        rc0, A0, j0, t0 = _call(handle, "hqrrp_geqp3_dev", A)
        rc1, A1, j1, t1 = _call(handle, "hqrrp_geqp3_noalloc_dev", A)
        assert rc0 == 0 and rc1 == 0
        assert torch.equal(j0, j1)
        assert torch.equal(t0, t1)
        assert torch.equal(A0, A1)
        # and a valid factorisation
        import oracle

        res, orth = oracle.residuals(A, np.asfortranarray(A1.numpy().T),
                                     j1.numpy(), t1.numpy()[:min(m, n)])
        eps = np.finfo(np.float64).eps
        assert res <= 50 * max(m, n) * eps and orth <= 50 * max(m, n) * eps


def test_noalloc_refuses_oversized(handle):
    import hqrrp_amd

    m, n = 2, 3_000_000  # 16 n bytes of norms > the scratch
    At = torch.ones((n, m), dtype=torch.float64, device="cuda")
    jp = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
    tau = torch.full((2,), 7.0, dtype=torch.float64, device="cuda")
    rc = hqrrp_amd.lib().hqrrp_geqp3_noalloc_dev(handle.ptr, m, n, _vp(At), m,
                                                 _vp(jp), _vp(tau))
    torch.cuda.synchronize()
    assert rc == -5
    assert torch.all(At == 1.0) and torch.all(tau == 7.0)
    assert torch.equal(jp, torch.arange(1, n + 1, dtype=torch.int32,
                                        device="cuda"))
