"""Shared helpers of the test-suite (budget closure, comparison)."""
from __future__ import annotations

import numpy as np

from sedcol.budget import ELEMENTS, closure as budget_residuals  # noqa: F401  (shared with bench.py)


def max_rel(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


def wet_cells(dom):
    off = dom.off
    return np.concatenate([np.arange(off[c], off[c + 1]) for c in range(dom.ncol) if not dom.mask[c]])


def rhs_terms(dom, params: dict, x, record=0):
    """(R, Q) of the oracle's reference-form right-hand side at state x:
    R the local reaction rates (the right-hand side with every transport
    path closed: no bioturbation, no molecular diffusion, closed top and
    bottom), Q = rhs - R the transport tendency."""
    import oracle as O

    p_ref = dict(params, arith=0)
    r, _ = O.Oracle(dom, p_ref, record=record).rhs(x)
    p_loc = dict(p_ref, bcup_dis=3, bcup_part=3, bcdown=3, D0=0.0)
    biot = dom.biot
    dom.biot = np.zeros_like(biot)
    try:
        R, _ = O.Oracle(dom, p_loc, record=record).rhs(x)
    finally:
        dom.biot = biot
    return R, r - R


def column_term_scale(dom, R, Q):
    """Per (variable, cell): the largest |R| + |Q| of the cell's column for
    that variable, the magnitude of the terms a right-hand side is made of."""
    s = np.abs(R) + np.abs(Q)
    off = dom.off
    out = np.empty_like(s)
    for c in range(dom.ncol):
        sl = slice(off[c], off[c + 1])
        out[:, sl] = s[:, sl].max(axis=1, keepdims=True) if off[c + 1] > off[c] else 0.0
    return out
