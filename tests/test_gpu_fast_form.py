"""The fast arithmetic form against the reference operation order on the
GPU, at the full size of the 300 x 300 shelf-sea domain (53 211 wet
columns, 1.6 M cells): size-independent bounds of the same kind as
test_fast_form.py asserts on the CPU oracle.
  * one right-hand side (sedcol_rhs) at the cold state and after one
    simulated day: |fast - reference| <= 16 ulp of the column's term scale
    max_k(|R_k| + |Q_k|) per variable, top fluxes within 4 ulp;
  * one simulated day of adaptive EMP (tolerance 1e-6) from the cold state:
    every component within 1e-10 (atol + |x|) of the reference-form run and
    within 10 rtol in tolerance units, every top flux within 1e-10
    relative, every concentration non-negative.
"""
import numpy as np
import pytest

from sedcol import cases, capi
from helpers import rhs_terms, column_term_scale

pytestmark = pytest.mark.gpu
ULP = 2.0 ** -52


@pytest.fixture(scope="module")
def shelf():
    dom = cases.shelf_sea(n=300, days=2, seed=0)
    states = {}
    for ar in (0, 1):
        dom.params.arith = ar
        with capi.SedimentColumns.from_domain(dom) as g:
            g.integrate(0.0, 86400.0)
            states[ar] = (g.get_state(), g.fluxes(), g.diagnostics())
    dom.params.arith = 0
    return dom, states


def _rhs_ulps(dom, x):
    out = {}
    for ar in (0, 1):
        dom.params.arith = ar
        with capi.SedimentColumns.from_domain(dom, record=0, series=False) as g:
            out[ar] = g.get_rhs(x)
    dom.params.arith = 0
    wet_cell = np.repeat(dom.mask == 0, dom.nlayers)
    R, Q = rhs_terms(dom, dom.params.si(), x)
    scale = column_term_scale(dom, R, Q)[:, wet_cell]
    err = np.abs(out[1][0] - out[0][0])[:, wet_cell]
    wet = dom.mask == 0
    f0, f1 = out[0][1][:, wet], out[1][1][:, wet]
    return float(np.max(err / np.maximum(scale, 1e-300))) / ULP, float(
        np.max(np.abs(f1 - f0) / np.maximum(np.abs(f0), 1e-300)))


@pytest.mark.parametrize("which", ["cold", "1 day"])
def test_fast_rhs_full_size(gpu_ready, shelf, which):
    dom, states = shelf
    x = dom.conc0.copy() if which == "cold" else states[0][0]
    ulps, ftop = _rhs_ulps(dom, x)
    print(f"full shelf-sea, {which}: max |fast - ref| = {ulps:.2f} ulp of the term scale, top flux {ftop:.1e}",
          flush=True)
    assert ulps <= 16.0
    assert ftop <= 4 * ULP


def test_fast_emp_day_full_size(gpu_ready, shelf):
    dom, states = shelf
    (x0, f0, d0), (x1, f1, d1) = states[0], states[1]
    wet = dom.mask == 0
    cells = np.repeat(wet, dom.nlayers)
    a, b = x0[:, cells], x1[:, cells]
    atol, rtol = dom.params.emp_atol, dom.params.emp_rtol
    dev = float(np.max(np.abs(b - a) / (atol + np.abs(a))))
    units = float(np.max(np.abs(b - a) / (atol + rtol * np.abs(a))))
    fdev = float(np.max(np.abs(f1 - f0)[:, wet] / np.maximum(np.abs(f0[:, wet]), 1e-300)))
    print(f"full shelf-sea EMP day: max |fast - ref|/(atol + |x|) = {dev:.2e} ({units:.2e} tolerance units), "
          f"top flux {fdev:.2e}, accepted {d0['accepted'].sum()} vs {d1['accepted'].sum()}", flush=True)
    assert np.all(a >= 0.0) and np.all(b >= 0.0)
    assert units <= 10.0
    assert dev <= 1e-10
    assert fdev <= 1e-10
