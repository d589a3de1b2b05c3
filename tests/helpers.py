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
