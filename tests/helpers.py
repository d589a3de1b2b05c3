"""Shared helpers of the test-suite (budget closure, comparison)."""
from __future__ import annotations

import numpy as np

ELEMENTS = {"C": ((0, 1), 11, 13), "N": ((2, 3, 6, 7), 12, 14), "P": ((4, 9, 10), None, 15)}


def max_rel(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


def budget_residuals(inv0, inv1, budget):
    """Relative closure of dI = F - sink + clamp per element and column.

    inv (NV, ncol); budget (ncol, 16).  The scale is the gross turnover
    |I0| + |I1| + |F| + |sink| + |clamp| of the element."""
    out = {}
    for e, (vs, isink, iclamp) in ELEMENTS.items():
        vs = list(vs)
        dI = inv1[vs].sum(0) - inv0[vs].sum(0)
        F = budget[:, vs].sum(1)
        sink = budget[:, isink] if isink is not None else 0.0
        cl = budget[:, iclamp]
        res = dI - (F - sink + cl)
        scale = np.abs(inv0[vs].sum(0)) + np.abs(inv1[vs].sum(0)) + np.abs(F) + np.abs(sink) + np.abs(cl)
        out[e] = np.abs(res) / np.maximum(scale, 1e-300)
    return out


def wet_cells(dom):
    off = dom.off
    return np.concatenate([np.arange(off[c], off[c + 1]) for c in range(dom.ncol) if not dom.mask[c]])
