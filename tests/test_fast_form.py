"""The fast arithmetic form (sedcol_params.arith = 1) against the reference
operation order (arith = 0) of get_rhs (fabm_sediment_driver.F90:652-715,
776-820), on the CPU oracle, which is bitwise equal to the GPU kernels in
either form (test_gpu_parity.py).

The fast form folds the geometry divisions into per-layer coefficients
(D/dzc, 1/(phi dz)), merges the Monod quotients and takes them through one
reciprocal: the same mathematics, different rounding.  Bounds asserted here:
  * one right-hand side: |fast - reference| <= 16 ulp of the column's term
    scale max_k(|R_k| + |Q_k|) per variable (R reaction, Q transport), on
    the cold and integrated states of the 4 x 4 RK4 case and of shelf-sea
    columns (measured: 5 ulp);
  * one RK4 step (dt = 360 s) from an integrated state: 1e-10 relative;
  * one simulated day of adaptive EMP (tolerance 1e-6) on shelf-sea columns
    from the cold state: every component within 1e-10 (atol + |x|) and every
    top flux within 1e-10 relative, i.e. far inside the 10 rtol that the
    step-size control allows (measured: 5e-14).
"""
import numpy as np
import pytest

from sedcol import cases
from sedcol.params import RUNGE_KUTTA_4
import oracle as O
from helpers import rhs_terms, column_term_scale

ULP = 2.0 ** -52


def _rhs_bound(dom, x, record=0):
    p = dom.params.si()
    r0, f0 = O.Oracle(dom, dict(p, arith=0), record=record).rhs(x)
    r1, f1 = O.Oracle(dom, dict(p, arith=1), record=record).rhs(x)
    R, Q = rhs_terms(dom, p, x, record)
    wet = np.repeat(dom.mask == 0, dom.nlayers)
    scale = column_term_scale(dom, R, Q)[:, wet]
    err = np.abs(r1 - r0)[:, wet]
    worst = float(np.max(err / np.maximum(scale, 1e-300)))
    ftop = float(np.max(np.abs(f1 - f0)[:, dom.mask == 0] / np.maximum(np.abs(f0[:, dom.mask == 0]), 1e-300)))
    return worst / ULP, ftop


@pytest.fixture(scope="module")
def rk4_states():
    d = cases.plain_columns(seed=0)
    x10 = d.conc0.copy()
    O.Oracle(d, d.params.si()).run(x10, 360.0, 2400, RUNGE_KUTTA_4, nthreads=4)
    return d, d.conc0.copy(), x10


@pytest.fixture(scope="module")
def shelf_states():
    s = cases.shelf_sea(n=60, days=2, seed=0)
    sub = s.subset(np.flatnonzero(s.mask == 0)[:48])
    x1 = sub.conc0.copy()
    O.Oracle(sub, dict(sub.params.si(), clamp=0)).emp_integrate(x1, 86400.0, nthreads=8)
    return sub, sub.conc0.copy(), x1


@pytest.mark.parametrize("which", ["cold", "10 days"])
def test_fast_rhs_within_ulps_rk4_case(rk4_states, which):
    d, x0, x10 = rk4_states
    ulps, ftop = _rhs_bound(d, x0 if which == "cold" else x10)
    print(f"4x4 RK4 case, {which}: max |fast - ref| = {ulps:.2f} ulp of the term scale, top flux {ftop:.1e}")
    assert ulps <= 16.0
    assert ftop <= 4 * ULP


@pytest.mark.parametrize("which", ["cold", "1 day"])
def test_fast_rhs_within_ulps_shelf_sea(shelf_states, which):
    d, x0, x1 = shelf_states
    ulps, ftop = _rhs_bound(d, x0 if which == "cold" else x1)
    print(f"shelf-sea columns, {which}: max |fast - ref| = {ulps:.2f} ulp of the term scale, top flux {ftop:.1e}")
    assert ulps <= 16.0
    assert ftop <= 4 * ULP


def test_fast_rk4_step_within_1e10(rk4_states):
    d, _, x10 = rk4_states
    p = d.params.si()
    out = []
    for ar in (0, 1):
        x = x10.copy()
        assert O.Oracle(d, dict(p, arith=ar)).ode_solver(x, 360.0, RUNGE_KUTTA_4) == 0
        out.append(x)
    rel = np.abs(out[1] - out[0]) / np.maximum(np.abs(out[0]), 1e-6)
    print(f"one RK4 step: max rel {rel.max():.2e}")
    assert rel.max() <= 1e-10


def test_fast_emp_day_within_tolerance():
    s = cases.shelf_sea(n=60, days=2, seed=0)
    sub = s.subset(np.flatnonzero(s.mask == 0)[::37][:16])
    p = dict(sub.params.si(), clamp=0)
    res = []
    for ar in (0, 1):
        x = sub.conc0.copy()
        orc = O.Oracle(sub, dict(p, arith=ar))
        assert orc.emp_integrate(x, 86400.0, nthreads=8) == 0
        res.append((x, orc.flux.copy(), orc.diag.copy()))
    (x0, f0, d0), (x1, f1, d1) = res
    atol, rtol = p["emp_atol"], p["emp_rtol"]
    dev = np.max(np.abs(x1 - x0) / (atol + np.abs(x0)))
    tol_units = np.max(np.abs(x1 - x0) / (atol + rtol * np.abs(x0)))
    fdev = np.max(np.abs(f1 - f0) / np.maximum(np.abs(f0), 1e-300))
    print(f"one EMP day: max |fast - ref|/(atol + |x|) = {dev:.2e}, in tolerance units {tol_units:.2e}, "
          f"top flux {fdev:.2e}; accepted {d0['accepted'].sum()} vs {d1['accepted'].sum()}")
    assert tol_units <= 10.0
    assert dev <= 1e-10
    assert fdev <= 1e-10
