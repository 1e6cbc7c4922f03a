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


def check_pivots(jp_gpu, jp_ref, trace_sketch, trace_panel, nb,
                 sketch_gpu=None, sketch_ref=None, panel_pivoting=1,
                 label=""):
    """Assert equal pivots, or a first divergence at a documented near-tie.

    Returns the index of the first divergence (None when identical) and
    prints every divergence with the step it is attributed to.

    The divergence at position i lies in block j = nb * (i // nb).  It is
    attributed to the exact decision that made it:

    * to block j's sketch when the block's set of columns differs (or the
      panel does not pivot): the first step s at which the GPU's sketch
      order (sketch_gpu, the library's trace) differs from the oracle's
      (sketch_ref); the oracle's gap between the two largest candidate
      sketch norms at step s must be below NEAR_TIE of the sketch's scale;
    * otherwise to the panel decision at position i, with the oracle's
      panel gap at i.

    Where the GPU's sketch order is not available (the batched kernel), the
    sketch step is bounded: the first oracle step t whose column is not in
    the GPU's block, the GPU having diverged at some step j..t; the smallest
    oracle gap over j..t must then be below NEAR_TIE.
    """
    jp_gpu = np.asarray(jp_gpu)
    jp_ref = np.asarray(jp_ref)
    i = first_divergence(jp_gpu, jp_ref)
    if i is None:
        return None
    j = (i // nb) * nb
    e = min(j + nb, len(jp_ref))
    blk_g = set(jp_gpu[j:e].tolist())
    blk_r = set(jp_ref[j:e].tolist())
    if blk_g != blk_r or not panel_pivoting:
        scale = trace_sketch[0][0]
        if sketch_gpu is not None:
            sg = np.asarray(sketch_gpu)[j:e]
            sr = np.asarray(sketch_ref)[j:e]
            d = np.nonzero(sg != sr)[0]
            assert d.size, (f"{label}: block {j} differs but its sketch "
                            "orders agree")
            s = j + int(d[0])
            g = rel_gap(trace_sketch[s], scale)
            where = f"sketch step {s}"
        else:
            sr = np.asarray(sketch_ref)[j:e]
            out = [k for k in range(e - j) if int(sr[k]) not in blk_g]
            t = j + (out[0] if out else e - j - 1)
            g = min(rel_gap(trace_sketch[k], scale) for k in range(j, t + 1))
            where = f"sketch steps {j}..{t}"
    else:
        g = rel_gap(trace_panel[i], trace_panel[0][0])
        where = f"panel step {i}"
    print(f"{label}: pivots diverge at {i} ({where}), oracle gap {g:.3e} "
          f"of the scale", flush=True)
    assert g < NEAR_TIE, (
        f"{label}: pivots diverge at {i} ({where}) where the oracle's gap is "
        f"{g:.3e} of the scale (>= {NEAR_TIE})")
    return i


def factor_residuals(A0, Afac, jpvt, tau):
    import oracle

    return oracle.residuals(A0, Afac, jpvt, tau)
