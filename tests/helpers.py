"""Comparison helpers shared by the parity tests."""
from __future__ import annotations

import numpy as np

# A pivot decision is a near-tie when the gap between the two largest
# candidate partial norms is below NEAR_TIE times the scale of the problem:
# the largest candidate norm of the first decision of the same kind (sketch or
# panel).  Deep in a graded spectrum the candidates are ~1e-11 of that scale,
# so a gap that is large relative to the current norms can still be pure
# rounding noise relative to ||A||; measuring against the initial scale is
# what separates a genuine disagreement from noise.
NEAR_TIE = 1e-12


def swaps_to_perm(piv, ncols):
    """Positions after applying swap(k, piv[k]) for k = 0.. to the identity."""
    p = np.arange(ncols)
    for k, q in enumerate(piv):
        p[k], p[q] = p[q], p[k]
    return p


def moves_to_perm(moves, ncols):
    p = np.arange(ncols)
    cnt = int(moves[0])
    for e in range(cnt):
        dst, src = int(moves[1 + 2 * e]), int(moves[2 + 2 * e])
        p[dst] = src
    return p


def first_divergence(jp_a, jp_b):
    d = np.nonzero(np.asarray(jp_a) != np.asarray(jp_b))[0]
    return None if d.size == 0 else int(d[0])


def rel_gap(tr_row, scale):
    best, second = tr_row
    return (best - second) / scale if scale > 0 else 0.0


def check_pivots(jp_gpu, jp_ref, trace_sketch, trace_panel, nb):
    """Assert equal pivots, or a first divergence at a documented near-tie.

    Returns the index of the first divergence (None when identical).  A
    divergence in block j is attributed to the sketch when the block's set of
    columns differs, otherwise to the panel ordering; the oracle's relative
    gap between the two largest candidates at the responsible steps must be
    below NEAR_TIE.
    """
    i = first_divergence(jp_gpu, jp_ref)
    if i is None:
        return None
    j = (i // nb) * nb
    blk_g = set(np.asarray(jp_gpu)[j:j + nb].tolist())
    blk_r = set(np.asarray(jp_ref)[j:j + nb].tolist())
    if blk_g != blk_r:
        scale = trace_sketch[0][0]
        gaps = [rel_gap(trace_sketch[k], scale)
                for k in range(j, min(j + nb, len(trace_sketch)))]
        where = "sketch"
    else:
        gaps = [rel_gap(trace_panel[i], trace_panel[0][0])]
        where = "panel"
    g = min(gaps)
    assert g < NEAR_TIE, (
        f"pivots diverge at {i} ({where} decision of block {j}) where the "
        f"oracle's smallest gap is {g:.3e} of the scale (>= {NEAR_TIE})")
    return i


def factor_residuals(A0, Afac, jpvt, tau):
    import oracle

    return oracle.residuals(A0, Afac, jpvt, tau)
