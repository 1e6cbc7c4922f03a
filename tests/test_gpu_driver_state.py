"""The driver's internal state read back through hqrrp_trace.

* Initial sketch (NoFLA_HQRRP_WY_blk_var4.c:407-409): Y = Omega A against
  the product formed in FP64 on the host from the reference's Omega, to
  FP64 rounding (FP32 and BF16 sketches to their own precision).
* The sketch downdate the driver runs (folded into the trailing update):
  after k blocks, the invariant the reference checks in its debug build
  (CHECK_DOWNDATING_OF_Y, NoFLA_HQRRP_WY_blk_var4.c:457-490),
  Y(:, k:) = G(:, k:) A(k:, k:), on the library's own Y, G' and trailing
  matrix; with and without look-ahead (bitwise equal to each other).
* Truncation (hqrrp_options.max_rank): the first k columns, the trailing
  matrix and the whole of jpvt against the oracle truncated the same way,
  including a cut that is not a multiple of nb.
"""
import numpy as np
import pytest

import oracle
from hqrrp_amd import matgen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = 2.0 ** -52


def _run(handle, A, nb, pp, trace_keys=(), sketch_prec=0, **kw):
    import hqrrp_amd

    m, n = A.shape
    r = nb + pp
    At = hqrrp_amd.to_colmajor(A)
    jp = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
    tau = torch.zeros(max(1, min(m, n)), dtype=torch.float64, device="cuda")
    ydt = torch.float64 if sketch_prec == 0 else torch.float32
    bufs = {}
    for key in trace_keys:
        if key in ("y0", "y"):
            bufs[key] = torch.full((n, r), float("nan"), dtype=ydt,
                                   device="cuda")
        elif key == "g":
            bufs[key] = torch.full((r, m), float("nan"), dtype=torch.float64,
                                   device="cuda")
        else:
            bufs[key] = torch.zeros(min(m, n), dtype=torch.int32,
                                    device="cuda")
    hqrrp_amd.hqrrp_device(handle, At, jp, tau, m, n, nb=nb, pp=pp,
                           sketch_prec=sketch_prec, trace=bufs, **kw)
    handle.synchronize()
    torch.cuda.synchronize()
    out = {k: v.double().cpu().numpy().T for k, v in bufs.items()
           if k != "sketch_cols"}  # y: r x n, g: m x r
    return (hqrrp_amd.from_colmajor(At), jp.cpu().numpy(),
            tau.cpu().numpy()[:min(m, n)], out)


def effective_rank(m, n, nb, kmax):
    """Columns a run truncated at kmax factors: blocks of nb up to kmax, the
    reference's final block (j + nb >= m or n) always whole."""
    k = min(m, n)
    for j in range(0, k, nb):
        if j + nb >= m or j + nb >= n:
            return k
        if j + nb >= kmax:
            return kmax
    return k


# Rounding bounds of Y = Omega A per column, relative to ||Omega||_2 ||A_c||,
# where Omega is the operator the sketch uses (the reference's, rounded to
# FP32 or bf16 values for a mixed sketch): a sum of m products in FP64 or
# FP32 (typically sqrt(m) eps; 8 sqrt(m) eps allowed), and for the BF16
# sketch A's two-term split (16 significant bits: 2^-17 relative per entry;
# 2^-14 allowed).
def _tol(prec, m):
    if prec == 0:
        return 8 * np.sqrt(m) * EPS
    if prec == 1:
        return 8 * np.sqrt(m) * 2.0 ** -24
    return 2.0 ** -14


def _round_operator(G, prec):
    """Omega rounded as the mixed sketch rounds it (to FP32, or to bf16 by
    round-to-nearest-even on the FP32 bits)."""
    if prec == 0:
        return G
    g = G.astype(np.float32)
    if prec == 2:
        b = g.view(np.uint32).astype(np.uint64)
        b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
        g = b.astype(np.uint32).view(np.float32)
    return g.astype(np.float64)


@pytest.mark.parametrize("m,n,prec", [(3000, 2000, 0), (1000, 3000, 0),
                                      (2000, 1500, 1), (2000, 1500, 2)])
def test_initial_sketch(handle, m, n, prec):
    """Y = Omega A against the FP64 product of the reference's Omega."""
    nb, pp = 64, 10
    print(f"initial sketch {m}x{n} prec={prec}", flush=True)
    A = matgen.gaussian(m, n, seed=11)
    G = _round_operator(oracle.omega(nb + pp, m), prec)
    _, _, _, st = _run(handle, A, nb, pp, ("y0",), sketch_prec=prec,
                       max_rank=1)
    Y = st["y0"]
    Yref = G @ A
    scale = np.linalg.norm(G, 2) * np.linalg.norm(A, axis=0)
    err = (np.abs(Y - Yref).max(axis=0) / scale).max()
    tol = _tol(prec, m)
    print(f"max column error {err:.2e} of ||Omega|| ||A_c|| (tol {tol:.1e})",
          flush=True)
    assert err <= tol


@pytest.mark.parametrize("m,n,nb,pp,blocks,prec", [
    (3000, 2500, 64, 10, 10, 0), (2000, 3000, 32, 8, 12, 0),
    (2500, 2000, 64, 10, 6, 1)])
def test_downdate_invariant(handle, m, n, nb, pp, blocks, prec):
    """Y_R = G_R A_R after `blocks` blocks, on the driver's own state."""
    k = blocks * nb
    print(f"downdate invariant {m}x{n} nb={nb} after {blocks} blocks "
          f"prec={prec}", flush=True)
    A = matgen.graded(m, n, seed=12, decades=6.0)
    Af, jp, tau, st = _run(handle, A, nb, pp, ("y", "g"), sketch_prec=prec,
                           max_rank=k, lookahead=1)
    Af0, jp0, tau0, st0 = _run(handle, A, nb, pp, ("y", "g"),
                               sketch_prec=prec, max_rank=k, lookahead=0)
    assert np.array_equal(jp, jp0) and np.array_equal(Af, Af0)
    assert np.array_equal(st["y"], st0["y"]) and np.array_equal(st["g"],
                                                                st0["g"])
    YR = st["y"][:, k:]
    GR = st["g"][k:, :].T
    AR = Af[k:, k:]
    err = np.linalg.norm(YR - GR @ AR)
    scale = np.linalg.norm(oracle.omega(nb + pp, m)) * np.linalg.norm(A)
    tol = 1e-13 if prec == 0 else 1e-6
    print(f"||Y_R - G_R A_R||_F = {err:.3e} = {err / scale:.2e} of "
          f"||Omega|| ||A||", flush=True)
    assert err <= tol * scale
    # A genuine downdate: the invariant of the un-downdated sketch fails.
    Y0 = oracle.omega(nb + pp, m) @ A[:, jp - 1]
    assert np.linalg.norm(Y0[:, k:] - GR @ AR) > 1e3 * tol * scale


@pytest.mark.parametrize("m,n,nb,pp,kmax,family", [
    (1500, 1200, 64, 10, 300, "gaussian"),     # cut block of 44 columns
    (1200, 1500, 32, 8, 256, "lowrank"),       # a multiple of nb
    (1000, 1000, 64, 10, 65, "gaussian"),      # cut block of 1 column
    (1000, 700, 64, 10, 690, "gaussian"),      # in the final block: whole
])
def test_truncated_matches_oracle(handle, m, n, nb, pp, kmax, family):
    print(f"truncated {family} {m}x{n} nb={nb} max_rank={kmax}", flush=True)
    if family == "gaussian":
        A = matgen.gaussian(m, n, seed=13)
    else:
        A = matgen.lowrank_noise(m, n, rank=400, seed=14)
    Ag, jg, tg, _ = _run(handle, A, nb, pp, max_rank=kmax)
    Ar, jr, tr = oracle.hqrrp(A, nb, pp, 1, max_cols=kmax)
    assert np.array_equal(jg, jr)
    # Columns done: R rows and reflectors; the rest: the trailing matrix.
    kk = effective_rank(m, n, nb, kmax)
    assert np.abs(tg[:kk] - tr[:kk]).max() <= 1e-12
    scale = max(1.0, np.abs(A).max())
    assert np.abs(Ag - Ar).max() <= 1e-11 * scale
    if kk < min(m, n):
        # tau beyond the cut is untouched (zero on entry)
        assert not np.any(tg[kk:])
