"""Pins of the oracle's linearly implicit stepper (method 7,
oracle/sedcol_oracle.c col_ros2_adaptive), a project definition with no
reference counterpart (the reference's ode_solver has methods 0-3 only,
solver_library.F90:80-189).

  * the analytic Jacobian of the local rates against central finite
    differences of oc_kinetics, and its C and N sink gradients against
    finite differences of Cprod and 0.8 Denit;
  * the block solve (I - g J_R) w = z against its definition;
  * the closed form of one step on a linear decay (no transport, exact
    Jacobian): c+ = c (1 + (1 - 2 gamma) z) / (1 - gamma z)^2, z = -k h,
    the stability function of ROS2 with gamma = 1 + 1/sqrt(2);
  * observed order 2 with the approximate (factored) W on columns with
    transport, where W differs from I - gamma h J;
  * positivity and the C/N/P budget closure (1e-12) on stiff cold starts and
    on spun-up columns, including the Dirichlet top-flux bookkeeping;
  * agreement with an RK4 reference at dt = 1 s within the tolerance.
A planted mistake in any of these (a sign, a coefficient, a dropped
boundary term) fails one of them.
"""
import math

import numpy as np
import pytest

from sedcol import cases, spinup
from sedcol.params import ROS2_ADAPTIVE, EMP_ADAPTIVE, RUNGE_KUTTA_4
import oracle as O
from helpers import budget_residuals

GAMMA = 1.0 + 1.0 / math.sqrt(2.0)


def _states(n=40, seed=3):
    rng = np.random.default_rng(seed)
    scale = np.array([5e4, 5e4, 8e3, 8e3, 500, 300, 40, 15, 200, 5, 20])
    out = [rng.uniform(0.0, 1.0, 11) * scale for _ in range(n)]
    out[0][5] = 0.0            # anoxic
    out[1][[5, 6]] = 0.0       # no O2, no NO3
    out[2][5] = 2.0            # near the ODU-oxidation half-saturation
    return out


def _fd_jac(p, fT, c):
    J = np.zeros((11, 11))
    gC = np.zeros(11)
    gN = np.zeros(11)
    for j in range(11):
        h = 1e-6 * max(abs(c[j]), 1.0)
        if c[j] - h < 0.0:
            # at the clamp: second-order one-sided (-3 f(c) + 4 f(c+h) - f(c+2h)) / 2h
            f = []
            for m in (0, 1, 2):
                a = c.copy()
                a[j] += m * h
                R, sC, sN = O.kinetics(p, fT, a)
                f.append(np.concatenate([R, [sC, sN]]))
            col = (-3.0 * f[0] + 4.0 * f[1] - f[2]) / (2.0 * h)
        else:
            a, b = c.copy(), c.copy()
            a[j] += h
            b[j] -= h
            Ra, sCa, sNa = O.kinetics(p, fT, a)
            Rb, sCb, sNb = O.kinetics(p, fT, b)
            col = (np.concatenate([Ra, [sCa, sNa]]) - np.concatenate([Rb, [sCb, sNb]])) / (a[j] - b[j])
        J[:, j], gC[j], gN[j] = col[:11], col[11], col[12]
    return J, gC, gN


def test_jacobian_against_finite_differences():
    p = cases.SedParams().si()
    for i, c in enumerate(_states()):
        fT = 0.7 + 0.02 * i
        Jfd, gCfd, gNfd = _fd_jac(p, fT, c)
        J = np.zeros((11, 11))
        gC = np.zeros(11)
        gN = np.zeros(11)
        for j in range(11):
            e = np.zeros(11)
            e[j] = 1.0
            J[:, j], _, gC[j], gN[j] = O.kin_jac_ops(p, fT, c, 1.0, e)
        # rounding noise of the differences: a few ulp of the rates over the step
        R, sC, sN = O.kinetics(p, fT, c)
        hstep = 1e-6 * np.maximum(np.abs(c), 1.0)
        noise = 64 * 2.2e-16 * (np.abs(R)[:, None] + sC) / hstep[None, :]
        scale = np.abs(Jfd).max(axis=1, keepdims=True) + 1e-300
        assert np.all(np.abs(J - Jfd) <= 1e-6 * scale + noise), (i, np.max(np.abs(J - Jfd) / scale))
        nC = 64 * 2.2e-16 * sC / hstep
        nN = 64 * 2.2e-16 * (sN + sC) / hstep
        assert np.all(np.abs(gC - gCfd) <= 1e-6 * (np.abs(gCfd).max() + 1e-300) + nC)
        assert np.all(np.abs(gN - gNfd) <= 1e-6 * (np.abs(gNfd).max() + 1e-300) + nN)


def test_local_solve_inverts_its_matrix():
    p = cases.SedParams().si()
    rng = np.random.default_rng(1)
    for c in _states():
        for g in (1.0, 300.0, 3.0e4):
            z = rng.standard_normal(11) * 100.0
            _, w, _, _ = O.kin_jac_ops(p, 1.0, c, g, z)
            assert w is not None
            Jw, _, _, _ = O.kin_jac_ops(p, 1.0, c, g, w)
            back = w - g * Jw
            assert np.max(np.abs(back - z)) <= 1e-12 * (np.abs(z).max() + g * np.abs(Jw).max())


def _decay_domain():
    d = cases.plain_columns(ni=2, nj=1, K=1, seed=3, total=0.001)
    d.biot[:] = 0.0
    d.fup[:] = 0.0
    d.cbw[:] = 0.0
    d.conc0[:] = 0.0
    d.conc0[0] = [100.0, 400.0]
    d.params.kin.CprodMax = 1.0e300
    d.params.D0 = 0.0
    d.params.bcup_dis = 3
    d.params.clamp = 0
    return d


@pytest.mark.parametrize("h", [60.0, 3600.0, 86400.0, 30 * 86400.0, 365 * 86400.0])
def test_one_step_linear_decay_closed_form(h):
    d = _decay_domain()
    p = dict(d.params.si(), adaptive_method=ROS2_ADAPTIVE, emp_dt_max=h, emp_rtol=1e3, emp_atol=1e3)
    orc = O.Oracle(d, p)
    x = d.conc0.copy()
    assert orc.emp_integrate(x, h) == 0
    assert np.all(orc.diag["accepted"] == 1)
    fT = np.array([O.oc_exp(-4500.0 * (1.0 / (t + 273.0) - 1.0 / 288.0)) for t in d.temp[0]])
    z = -fT * p["rL"] * h
    R = (1.0 + (1.0 - 2.0 * GAMMA) * z) / (1.0 - GAMMA * z) ** 2
    rel = np.abs(x[0] - d.conc0[0] * R) / (d.conc0[0] * R)
    print(f"h = {h}: z = {z}, max rel {rel.max():.2e}")
    assert rel.max() <= 1e-13


def test_observed_order_two_with_the_factored_w():
    """A linear problem with transport: labile and semi-labile detritus decay
    (production cap 1, anoxic, so all carbon goes to ODU) while bioturbation
    and diffusion mix 8 layers.  Fixed steps (loose tolerance, dt_max = h)
    against the exact solution expm(L T) c0: the factored W is not
    I - gamma h L, and the ROS2 W-method keeps order 2."""
    from scipy.linalg import expm

    d = cases.plain_columns(ni=1, nj=1, K=8, seed=3, total=0.02)
    d.fup[:] = 0.0
    d.cbw[:] = 0.0
    d.conc0[:] = 0.0
    z = np.linspace(0.0, 1.0, 8)
    d.conc0[0] = 1e4 * np.exp(-3 * z)
    d.conc0[1] = 5e3 * (1 + z)
    d.conc0[8] = 50.0 * z
    d.biot[:] = 20.0
    d.params.kin.CprodMax = 1.0e300
    d.params.kin.rLabile = 2.0
    d.params.kin.rSemilabile = 0.5
    d.params.clamp = 0
    p = d.params.si()
    orc = O.Oracle(d, p)
    N = 11 * 8
    L = np.zeros((N, N))
    for j in range(N):
        e = np.zeros((11, 8))
        e.flat[j] = 1.0
        L[:, j] = orc.rhs(e)[0].ravel()
    T = 6 * 3600.0
    exact = (expm(L * T) @ d.conc0.ravel()).reshape(11, 8)
    errs = []
    for n in (32, 64, 128):
        x = d.conc0.copy()
        o = O.Oracle(d, dict(p, adaptive_method=ROS2_ADAPTIVE, emp_dt_max=T / n, emp_rtol=1e6, emp_atol=1e6))
        assert o.emp_integrate(x, T) == 0
        assert o.diag["accepted"][0] == n
        errs.append(np.max(np.abs(x - exact)) / np.abs(exact).max())
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print("errors", errs, "orders", orders)
    assert 1.8 < orders[0] < 2.2 and 1.9 < orders[1] < 2.1, orders


@pytest.mark.parametrize("which", ["cold", "spun-up"])
def test_positivity_and_budget_closure(which):
    rng = np.random.Generator(np.random.Philox(5))
    s = spinup.draw_inputs(rng, 6)
    d = spinup.build(s, np.array([30, 30, 22, 40, 17, 30]))
    p = dict(d.params.si(), clamp=0)
    x = d.conc0.copy()
    if which == "spun-up":
        O.Oracle(d, p).emp_integrate(x, 30 * 86400.0, nthreads=6)
    orc = O.Oracle(d, dict(p, adaptive_method=ROS2_ADAPTIVE))
    inv0 = orc.inventory(x)
    assert orc.emp_integrate(x, 86400.0, nthreads=6) == 0
    assert np.all(x >= 0.0)
    res = budget_residuals(inv0, orc.inventory(x), orc.budget())
    for e, r in res.items():
        print(which, e, r.max())
        assert r.max() <= 1e-12, (e, r.max())


def test_shelf_sea_records_budget_and_dirichlet_fluxes():
    """Daily forcing records on shelf-sea columns (Dirichlet solute tops)."""
    full = cases.shelf_sea(n=30, days=3, seed=5)
    dom = full.subset(np.flatnonzero(full.mask == 0)[:16])
    p = dict(dom.params.si(), adaptive_method=ROS2_ADAPTIVE)
    orc = O.Oracle(dom, p)
    x = dom.conc0.copy()
    spans = cases.record_spans(0.0, 3 * 86400.0, dom.nrec, dom.t0, dom.dt_rec)
    assert orc.integrate_series(x, 0.0, 3 * 86400.0, spans, nthreads=4) == 0
    assert np.all(x >= 0.0)
    res = budget_residuals(orc.inventory(dom.conc0), orc.inventory(x), orc.budget())
    for e, r in res.items():
        assert r.max() <= 1e-12, (e, r.max())


def test_accuracy_against_rk4():
    """One spun-up column over 2 hours: ROS2 at tolerance 1e-6 against RK4
    at dt = 1 s (whose own error is far smaller)."""
    rng = np.random.Generator(np.random.Philox(5))
    s = spinup.draw_inputs(rng, 2)
    d = spinup.build(s, np.full(2, 30))
    p = dict(d.params.si(), clamp=0)
    x0 = d.conc0.copy()
    O.Oracle(d, p).emp_integrate(x0, 20 * 86400.0)
    T = 2 * 3600.0
    ref = x0.copy()
    O.Oracle(d, p).run(ref, 1.0, int(T), RUNGE_KUTTA_4, nthreads=2)
    x = x0.copy()
    orc = O.Oracle(d, dict(p, adaptive_method=ROS2_ADAPTIVE))
    assert orc.emp_integrate(x, T) == 0
    dev = np.max(np.abs(x - ref) / (p["emp_atol"] + np.abs(ref)))
    print(f"ROS2 vs RK4(1 s): {dev:.2e} relative, accepted {orc.diag['accepted']}")
    assert dev <= 1e-4
