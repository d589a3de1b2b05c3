"""C/N/P budget closure of sediment columns from the library's inventories
and accumulated budgets (sedcol_get_inventory, sedcol_get_budget).

Per element and column: inventory change = integrated top fluxes - released
DIC (C) or N2 (N) + clamp mass.  Budget columns (sedcol.h): 0..10 the
time-integrated top flux of each variable, 11 DIC, 12 N2, 13..15 clamped C,
N, P."""
from __future__ import annotations

import numpy as np

#: element -> (state variables carrying it, sink column or None, clamp column)
ELEMENTS = {"C": ((0, 1), 11, 13), "N": ((2, 3, 6, 7), 12, 14), "P": ((4, 9, 10), None, 15)}


def closure(inv0, inv1, budget):
    """Relative closure per element and column: |dI - (F - sink + clamp)|
    over the gross turnover |I0| + |I1| + |F| + |sink| + |clamp|.

    inv0, inv1: (NV, ncol) inventories; budget: (ncol, 16)."""
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
