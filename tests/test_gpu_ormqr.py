"""hqrrp_ormqr_dev (LAPACK dormqr semantics) against the oracle's dormqr
(LAPACK, pinned on the CPU against the explicit Q of dorgqr): all four
side/trans combinations, k below, at and above one 64-column block and not
a multiple of it, reflectors from the library's own HQRRP."""
import numpy as np
import pytest

import oracle
from hqrrp_amd import matgen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("side,trans", [("L", "N"), ("L", "T"), ("R", "N"),
                                        ("R", "T")])
@pytest.mark.parametrize("nq,kcols,k,other", [(700, 300, 200, 150),
                                              (500, 64, 64, 90),
                                              (1300, 1000, 1000, 33)])
def test_ormqr_matches_lapack(handle, side, trans, nq, kcols, k, other):
    import hqrrp_amd

    print(f"ormqr {side}{trans} nq={nq} k={k}", flush=True)
    A = matgen.gaussian(nq, kcols, 72)
    Af, _, tau = hqrrp_amd.hqrrp(A, 64, 10, 1)
    tau = tau[:k]
    m, n = (nq, other) if side == "L" else (other, nq)
    Cm = np.asfortranarray(np.random.default_rng(73).standard_normal((m, n)))
    want = oracle.dormqr(side, trans, Af, tau, Cm)
    At = hqrrp_amd.to_colmajor(Af)
    Ct = hqrrp_amd.to_colmajor(Cm)
    tt = torch.as_tensor(np.ascontiguousarray(tau)).cuda()
    hqrrp_amd.ormqr_device(handle, side, trans, At, tt, k, Ct, m, n)
    handle.synchronize()
    torch.cuda.synchronize()
    got = hqrrp_amd.from_colmajor(Ct)
    err = np.abs(got - want).max() / np.abs(Cm).max()
    print(f"max difference {err:.2e} of max|C|", flush=True)
    assert err <= 1e-13


def test_ormqr_argument_errors(handle):
    import hqrrp_amd

    L = hqrrp_amd.lib()
    z = hqrrp_amd.api._vp(torch.zeros(16, dtype=torch.float64, device="cuda"))
    assert L.hqrrp_ormqr_dev(handle.ptr, b"X", b"N", 4, 4, 2, z, 4, z, z,
                             4) == -1
    assert L.hqrrp_ormqr_dev(handle.ptr, b"L", b"Q", 4, 4, 2, z, 4, z, z,
                             4) == -1
    assert L.hqrrp_ormqr_dev(handle.ptr, b"L", b"N", 4, 4, 5, z, 4, z, z,
                             4) == -1  # k > m
    assert L.hqrrp_ormqr_dev(handle.ptr, b"R", b"T", 4, 3, 2, z, 2, z, z,
                             4) == -1  # lda < n
    assert L.hqrrp_ormqr_dev(handle.ptr, b"L", b"N", 0, 4, 0, z, 1, z, z,
                             1) == 0
