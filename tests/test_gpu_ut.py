"""The libflame UT-variant front end (hqrrp_ut_blk_var2, hqrrp_ut_factor_dev;
FLA_HQRRP_UT_blk_var2.c:124-401) against the oracle's restatement of that
driver (libflame's UT Householder included): the same pivots as the oracle
(near-ties as in tests/test_gpu_parity.py), s and the factor to rounding
where the reflectors are well determined, and the UT transform's identity
A P = Q R with Q = I - U inv(T) U' on the library's own output.  libflame is
not available, so the oracle's UT driver is itself pinned on the CPU (by the
same identity and by the WY driver, tests/test_oracle.py)."""
import numpy as np
import pytest

import oracle
from hqrrp_amd import matgen
from tests.helpers import check_pivots
from tests.test_oracle import _libflame_simple_test_matrix

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = np.finfo(np.float64).eps

CASES = [
    # (kind, m, n, nb, pp, seed)
    ("libflame_simple_test", 5, 5, 64, 10, 0),
    ("gaussian", 1200, 1000, 64, 10, 111),
    ("gaussian", 900, 1300, 64, 10, 112),   # wide: the last step reflects
    ("graded", 1500, 1500, 64, 10, 113),
    ("gaussian", 700, 600, 32, 8, 114),
]


def _matrix(kind, m, n, seed):
    if kind == "libflame_simple_test":
        return _libflame_simple_test_matrix()
    return getattr(matgen, kind)(m, n, seed)


def _identity(A, Af, jp, s):
    m, n = A.shape
    k = min(m, n)
    Q = oracle.ut_q(Af, s)
    res = np.linalg.norm(A[:, jp - 1] - Q[:, :k] @ np.triu(Af)[:k, :]) / \
        np.linalg.norm(A)
    return res, np.abs(Q.T @ Q - np.eye(m)).max()


@pytest.mark.parametrize("kind,m,n,nb,pp,seed", CASES)
def test_ut_front_end_matches_oracle(handle, kind, m, n, nb, pp, seed):
    import hqrrp_amd

    print(f"UT {kind} {m}x{n} nb={nb}", flush=True)
    A = _matrix(kind, m, n, seed)
    k = min(m, n)
    Ag, jg, sg = hqrrp_amd.hqrrp_ut(A, nb, pp, 1)
    Ar, jr, sr = oracle.hqrrp_ut(A, nb, pp, 1)
    _, _, _, ts, tp, _ = oracle.hqrrp(A, nb, pp, 1, trace=True)
    div = check_pivots(jg, jr, ts, tp, nb, label=f"UT {kind} {m}x{n}")
    assert div is None, f"pivots diverge at {div}"
    res, orth = _identity(A, Ag, jg, sg)
    print(f"UT identity {res:.2e} orthogonality {orth:.2e}", flush=True)
    assert res <= 50 * max(m, n) * EPS and orth <= 50 * max(m, n) * EPS
    # the reflections where nothing is left below the diagonal
    assert np.array_equal(sg == 0.5, sr == 0.5)
    dr = np.abs(np.diag(Ar)[:k])
    keep = dr >= 1e-5 * dr[0]
    assert (np.abs(np.abs(np.diag(Ag)[:k]) - dr)[keep] / dr[keep]).max() <= 1e-10
    assert (np.abs(sg - sr)[keep] / sr[keep]).max() <= 1e-10
    # R's signs follow libflame's convention, reflected rows included
    assert np.array_equal(np.sign(np.diag(Ag)[:k])[keep],
                          np.sign(np.diag(Ar)[:k])[keep])


def test_ut_device_entry_equals_host_entry(handle):
    """hqrrp_ut_factor_dev on device arrays: bitwise what the host entry
    returns; equal to hqrrp_factor_dev's result rewritten (s = 1/tau, the
    reflected rows negated) -- the same factorisation."""
    import hqrrp_amd

    m, n = 1100, 1300
    A = matgen.gaussian(m, n, 115)
    Ah, jh, sh = hqrrp_amd.hqrrp_ut(A, 64, 10, 1)
    At = hqrrp_amd.to_colmajor(A)
    jp = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
    s = torch.zeros(min(m, n), dtype=torch.float64, device="cuda")
    hqrrp_amd.hqrrp_ut_device(handle, At, jp, s, m, n)
    handle.synchronize()
    torch.cuda.synchronize()
    Ad = hqrrp_amd.from_colmajor(At)
    assert np.array_equal(Ad, Ah)
    assert np.array_equal(jp.cpu().numpy(), jh)
    assert np.array_equal(s.cpu().numpy(), sh)
    Aw, jw, tw = hqrrp_amd.hqrrp(A, 64, 10, 1)
    assert np.array_equal(jw, jh)
    deg = tw == 0.0
    assert deg[-1] and np.array_equal(sh[deg], np.full(deg.sum(), 0.5))
    assert np.array_equal(sh[~deg], 1.0 / tw[~deg])
    sign = np.where(deg, -1.0, 1.0)[:, None]
    k = min(m, n)
    assert np.array_equal(np.triu(Ah)[:k], sign * np.triu(Aw)[:k])
    assert np.array_equal(np.tril(Ah, -1), np.tril(Aw, -1))


def test_ut_device_entry_truncated(handle):
    """With hqrrp_options.max_rank the rewrite covers exactly the factored
    columns: s and R's reflected rows for k < kmax, the rest untouched
    (equal to the WY device entry's output there)."""
    import hqrrp_amd

    m, n, kmax = 1300, 1200, 300
    A = matgen.gaussian(m, n, 116)

    def run(ut):
        At = hqrrp_amd.to_colmajor(A)
        jp = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
        s = torch.zeros(min(m, n), dtype=torch.float64, device="cuda")
        f = hqrrp_amd.hqrrp_ut_device if ut else hqrrp_amd.hqrrp_device
        f(handle, At, jp, s, m, n, max_rank=kmax)
        handle.synchronize()
        torch.cuda.synchronize()
        return hqrrp_amd.from_colmajor(At), jp.cpu().numpy(), s.cpu().numpy()

    Au, ju, su = run(True)
    Aw, jw, tw = run(False)
    assert np.array_equal(ju, jw)
    assert np.array_equal(su[:kmax], 1.0 / tw[:kmax])  # no reflection here
    assert np.array_equal(su[kmax:], tw[kmax:])
    assert np.array_equal(Au, Aw)
