"""CPU tests that pin the oracle (oracle/sedcol_oracle.c).

What pins what:
  * steppers (solver_library.F90:80-189) -- the reference's own
    solver_library.F90 compiled into oracle/_ref (when built here) and the
    golden vectors it wrote (tests/golden/ref_steppers.npz), bitwise; closed
    forms for linear decay (RK4 and RK4-3/8 reproduce the degree-4 Taylor
    polynomial, Euler 1+z);
  * ADAPTIVE_EULER's global rule -- reference build / golden vectors, and the
    no-reject special case equal to EULER;
  * transport (get_rhs + diff3d) -- telescoping sum phi dz Q = F_top - F_bottom,
    the uniform Dirichlet steady state, zero divergence of a linear profile,
    the reference's layer-height / porosity formulas;
  * kinetics (library definition; FABM hzg_omexdia_p is not available, parity
    with it is unpinned) -- C/N/P stoichiometry, oxic/anoxic limits, zero
    destruction of an exhausted species;
  * EMP -- the closed form c/(1 + k dt) of one linear sink, positivity at any
    step, conservation, observed orders 1 and 2, for both Patankar weightings;
    the geometric-mean factor of both stages against a bracketing root
    solve of its defining equation;
  * exp/rsqrt -- against libm.
A planted mistake in any of these formulas fails at least one test here.
"""
import dataclasses
import math

import numpy as np
import pytest

from sedcol import cases, grid
from sedcol.params import EULER, RUNGE_KUTTA_4, ADAPTIVE_EULER, RUNGE_KUTTA_4_38, EMP1, EMP2
import oracle as O
from helpers import budget_residuals

GOLDEN = __import__("pathlib").Path(__file__).with_name("golden") / "ref_steppers.npz"


# ---------------------------------------------------------------------------
# reference steppers
# ---------------------------------------------------------------------------

def _golden_domain():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", GOLDEN.with_name("make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg


@pytest.mark.parametrize("name,method,dt,nsteps", [("euler", 0, 60.0, 20), ("rk4", 1, 360.0, 40),
                                                   ("adaptive_euler", 2, 3600.0, 4), ("rk4_38", 3, 360.0, 40)])
def test_oracle_matches_reference_golden(name, method, dt, nsteps):
    mg = _golden_domain()
    z = np.load(GOLDEN)
    d = mg.domain()
    orc = O.Oracle(d, d.params.si())
    x = z["x0"].copy()
    for _ in range(nsteps):
        assert orc.ode_solver(x, dt, method) == 0
    assert np.array_equal(x, z[name]), f"{name}: max abs diff {np.max(np.abs(x - z[name]))}"
    if method == ADAPTIVE_EULER:
        assert orc.last_min_dt[0] == z[name + "_last_min_dt"][0]
        assert orc.last_min_dt[0] < dt


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("method,dt,nsteps", [(EULER, 100.0, 15), (RUNGE_KUTTA_4, 720.0, 25),
                                              (ADAPTIVE_EULER, 7200.0, 3), (RUNGE_KUTTA_4_38, 720.0, 25)])
def test_oracle_matches_reference_build(method, dt, nsteps):
    d = cases.plain_columns(ni=3, nj=1, K=9, seed=8, total=0.08)
    d.params.dt_min = 10.0
    d.params.clamp = 0
    p = d.params.si()
    rng = np.random.default_rng(5)
    x0 = d.conc0.copy()
    x0[0:2] = rng.uniform(0.0, 3.0e4, size=x0[0:2].shape)
    x0[8] = rng.uniform(0.0, 50.0, size=x0[8].shape)
    a = x0.copy()
    ref = O.Oracle(d, p)
    ref.ref_ode_solver(a, dt, method, nsteps=nsteps)
    b = x0.copy()
    orc = O.Oracle(d, p)
    for _ in range(nsteps):
        orc.ode_solver(b, dt, method)
    assert np.array_equal(a, b)
    if method == ADAPTIVE_EULER:
        assert ref.last_min_dt[0] == orc.last_min_dt[0]


def _decay_domain(K=4, ncol=2):
    """Columns where only labile detritus is present and nothing moves:
    bioturbation 0, no deposition, CprodMax so large that the production cap
    is exactly 1, so dC/dt = -kL C exactly."""
    d = cases.plain_columns(ni=ncol, nj=1, K=K, seed=3, total=0.02)
    d.biot[:] = 0.0
    d.fup[:] = 0.0
    d.cbw[:] = 0.0
    d.conc0[:] = 0.0
    d.conc0[0] = np.linspace(100.0, 500.0, d.ncell)
    d.params.kin.CprodMax = 1.0e300
    d.params.clamp = 0
    return d


@pytest.mark.parametrize("method", [EULER, RUNGE_KUTTA_4, RUNGE_KUTTA_4_38])
def test_linear_decay_closed_form(method):
    d = _decay_domain()
    p = d.params.si()
    dt = 6.0 * 3600.0
    T = d.temp[0]
    fT = np.array([O.oc_exp(-4500.0 * (1.0 / (t + 273.0) - 1.0 / 288.0)) for t in T])
    k = fT * p["rL"]
    x = d.conc0.copy()
    orc = O.Oracle(d, p)
    n = 10
    for _ in range(n):
        orc.ode_solver(x, dt, method)
    z = -np.repeat(k, d.nlayers) * dt
    g = 1.0 + z if method == EULER else 1.0 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    expect = d.conc0[0] * g ** n
    assert np.allclose(x[0], expect, rtol=1e-13, atol=0)
    # the other particulate pools stay empty (the dissolved products diffuse
    # with explicit steps far above their stability limit and are not checked)
    assert np.all(x[1:5] == 0.0)


def test_adaptive_euler_without_rejects_is_euler():
    d = _decay_domain()
    p = d.params.si()
    a = d.conc0.copy()
    b = d.conc0.copy()
    O.Oracle(d, p).ode_solver(a, 600.0, ADAPTIVE_EULER)
    O.Oracle(d, p).ode_solver(b, 600.0, EULER)
    assert np.array_equal(a, b)


# ---------------------------------------------------------------------------
# transport
# ---------------------------------------------------------------------------

def _transport_only(d):
    """Zero every reaction: no detritus reactivity, no nitrification, ..."""
    k = d.params.kin
    k.rLabile = k.rSemilabile = k.rnit = k.rODUox = k.PAds = 0.0
    return d


@pytest.mark.parametrize("arith", [0, 1])
@pytest.mark.parametrize("bcup_dis,bcup_part", [(2, 1), (1, 1), (3, 3)])
def test_transport_telescopes(arith, bcup_dis, bcup_part):
    d = _transport_only(cases.plain_columns(ni=3, nj=2, K=13, seed=2))
    d.params.arith = arith
    d.params.bcup_dis, d.params.bcup_part = bcup_dis, bcup_part
    p = d.params.si()
    rng = np.random.default_rng(1)
    x = rng.uniform(0.0, 100.0, size=d.conc0.shape)
    r, f = O.Oracle(d, p).rhs(x)
    w = d.phi * d.dz
    off = d.off
    for c in range(d.ncol):
        sl = slice(off[c], off[c + 1])
        for n in range(11):
            s = np.sum(w[sl] * r[n, sl])
            scale = np.sum(np.abs(w[sl] * r[n, sl])) + abs(f[n, c]) + 1e-300
            assert abs(s - f[n, c]) <= 1e-13 * scale, (c, n, s, f[n, c])
    if bcup_dis == 3:
        assert np.all(f[5:10] == 0.0)


def test_dirichlet_uniform_profile_is_steady():
    d = _transport_only(cases.plain_columns(ni=2, nj=2, K=10, seed=4))
    d.fup[:] = 0.0
    no3 = d.cbw[:, 6].copy()
    d.cbw[:] = 0.0
    d.cbw[:, 6] = no3
    p = d.params.si()
    x = np.zeros_like(d.conc0)
    colid = np.repeat(np.arange(d.ncol), d.nlayers)
    x[6] = d.cbw[0, 6][colid]  # nitrate equal to bottom water everywhere
    r, f = O.Oracle(d, p).rhs(x)
    assert np.all(r == 0.0)
    assert np.all(f[6] == 0.0)


def test_linear_profile_zero_interior_divergence():
    d = _transport_only(cases.plain_columns(ni=1, nj=1, K=12, seed=4))
    K = 12
    d.dz[:] = 0.01
    d.dzc[1:] = 0.01
    d.phi[:] = 0.8
    d.phiI[:] = 0.8
    d.bfac[:] = 1.0
    p = d.params.si()
    x = np.zeros_like(d.conc0)
    x[6] = 5.0 + 2.0 * np.arange(K)
    r, _ = O.Oracle(d, p).rhs(x)
    assert np.allclose(r[6, 1:K - 1], 0.0, atol=1e-20)
    assert r[6, K - 1] < 0.0  # closed bottom: the deepest layer loses to the one above


def test_grid_formulas():
    g = grid.init_grid_linear(30, 0.001)
    assert math.isclose(g["dz"].sum(), 0.18 * 30 / 31, rel_tol=1e-13)
    assert math.isclose(g["dz"][0], 0.001, rel_tol=1e-15)
    g = grid.init_grid_geometric(30, 1.0e-3, 0.2)
    assert math.isclose(g["dz"].sum(), 0.2, rel_tol=1e-12)
    assert math.isclose(g["dz"][1] / g["dz"][0], 1.110256, rel_tol=1e-6)
    assert np.allclose(g["dzc"][1:], 0.5 * (g["dz"][1:] + g["dz"][:-1]), rtol=1e-13)
    phi = grid.porosity_linear(g["zc"])
    pi = grid.interface_porosity(phi)
    assert pi[0] == phi[0] and np.allclose(pi[1:], 0.5 * (phi[1:] + phi[:-1]))
    b = grid.bioturbation_factor(g["zi"][:-1], 1, bioturbation=0.9, depth_cm=5.0, min_cm2d=0.2)
    assert b[0] == 1.0 and np.isclose(b.min(), 0.2 / 0.9)


# ---------------------------------------------------------------------------
# kinetics (library definition)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("arith", [0, 1])
def test_kinetics_stoichiometry(arith):
    p = dict(cases.SedParams(arith=arith).si())
    rng = np.random.default_rng(9)
    for _ in range(200):
        c = rng.uniform(0.0, 1.0, 11) * np.array([5e4, 5e4, 8e3, 8e3, 500, 300, 40, 15, 200, 5, 20])
        fT = rng.uniform(0.5, 1.5)
        R, sC, sN = O.kinetics(p, fT, c)
        scale = np.abs(R).max() + 1e-300
        assert abs(R[0] + R[1] + sC) <= 4e-16 * scale * 4
        assert abs(R[2] + R[3] + R[6] + R[7] + sN) <= 1e-15 * scale * 8
        assert abs(R[4] + R[9] + R[10]) <= 1e-15 * scale * 8


def test_kinetics_limits():
    p = dict(cases.SedParams().si())
    c = np.array([1e4, 0, 0, 0, 0, 1e9, 0, 0, 0, 0, 0], dtype=float)
    R, sC, sN = O.kinetics(p, 1.0, c)
    assert math.isclose(-R[5], sC, rel_tol=1e-6)  # pure oxic mineralisation: O2 use = C production
    c = np.array([1e4, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0], dtype=float)
    R, sC, sN = O.kinetics(p, 1.0, c)
    assert R[5] == 0.0 and R[6] == 0.0 and math.isclose(R[8], sC, rel_tol=1e-15)  # all anoxic -> ODU
    assert sN == 0.0


def test_kinetics_exhausted_species_not_destroyed():
    p = dict(cases.SedParams().si())
    rng = np.random.default_rng(4)
    for j in range(11):
        c = rng.uniform(1.0, 100.0, 11)
        c[j] = 0.0
        R, _, _ = O.kinetics(p, 1.1, c)
        assert R[j] >= 0.0, j


# ---------------------------------------------------------------------------
# EMP
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("weights", [0, 1])
def test_emp1_single_linear_sink_closed_form(weights):
    d = _decay_domain(K=1, ncol=1)
    d.params.emp_weights = weights
    p = d.params.si()
    k = O.oc_exp(-4500.0 * (1.0 / (d.temp[0, 0] + 273.0) - 1.0 / 288.0)) * p["rL"]
    for dt in (10.0, 1.0e5, 1.0e9):
        x = d.conc0.copy()
        O.Oracle(d, p).ode_solver(x, dt, EMP1)
        # c + dt p f cancels down to c/(1 + k dt): relative rounding grows
        # with 1 + k dt (one destroyed component: both weightings solve the
        # same equation; the geometric solve goes through exp and log)
        tol = (4e-16 if weights == 0 else 1e-13) * (1.0 + k * dt)
        assert math.isclose(x[0, 0], d.conc0[0, 0] / (1.0 + k * dt), rel_tol=tol)


def _geo_factor(c, g, den, dt):
    """Root of p = (S prod (1 + a_j p))^(1/N) by bracketing (independent of
    the oracle's Newton iteration): a_j = dt g_j / c_j over g_j < 0, c_j > 0,
    S = prod c_j / den_j over those with den_j > 0 (second stage)."""
    m = (g < 0.0) & (c > 0.0)
    a = dt * g[m] / c[m]
    if a.size == 0:
        return 1.0
    lnS = 0.0
    if den is not None:
        dm = m & (den > 0.0)
        lnS = float(np.sum(np.log(c[dm] / den[dm])))
    N = a.size

    def f(p):  # decreasing on (0, 1/max|a|), +inf at 0 and -inf at the end
        with np.errstate(divide="ignore"):
            return lnS + float(np.sum(np.log1p(a * p))) - N * math.log(p)

    lo, hi = 1e-300, 1.0 / float(np.max(-a))
    for _ in range(400):
        mid = 0.5 * (lo + hi)
        if mid in (lo, hi):
            break
        if f(mid) > 0.0:
            lo = mid
        else:
            hi = mid
    return lo


@pytest.mark.parametrize("method", [EMP1, EMP2])
def test_emp_geometric_factor_against_root_solve(method):
    """One EMP step with geometric-mean weights equals the step built from
    bracketing roots of the factor equation (a per-column mean, not the
    product, and S only in the second stage).  The product weighting gives
    a different step, so a dropped 1/N fails here."""
    d = cases.plain_columns(ni=1, nj=1, K=12, seed=21)
    d.params.clamp = 0
    d.params.emp_weights = 1
    p = d.params.si()
    orc = O.Oracle(d, p)
    x0 = d.conc0.copy()
    rng = np.random.default_rng(5)
    x0[0:5] = rng.uniform(10.0, 1.0e3, size=x0[0:5].shape)
    x0[8] = rng.uniform(0.0, 50.0, size=x0[8].shape)
    x0[10] = rng.uniform(0.0, 10.0, size=x0[10].shape)
    dt = 1800.0
    f0, _ = orc.rhs(x0)
    p1 = _geo_factor(x0, f0, None, dt)
    c1 = np.maximum(x0 + (dt * p1) * f0, 0.0)
    if method == EMP1:
        expect = c1
    else:
        f1, _ = orc.rhs(c1)
        g = (f0 + f1) * 0.5
        p2 = _geo_factor(x0, g, c1, dt)
        expect = np.maximum(x0 + (dt * p2) * g, 0.0)
    x = x0.copy()
    assert orc.ode_solver(x, dt, method) == 0
    assert np.allclose(x, expect, rtol=1e-11, atol=1e-12)
    # the product weighting's step differs
    q = dataclasses.replace(d.params, emp_weights=0)
    y = x0.copy()
    assert O.Oracle(d, q.si()).ode_solver(y, dt, method) == 0
    assert not np.allclose(y, expect, rtol=1e-6, atol=1e-9)


def test_geometric_weights_hold_the_tolerance_on_a_long_column():
    """Cold start of a 60-layer column, 10 minutes, adaptive EMP at tolerance
    1e-6 with dt >= 1 s, against RK4 with dt = 0.01 s (which agrees with
    either EMP form at tolerance 1e-10 to 1e-8).  Both forms are held at
    dt_min by their error estimate, but the product factor throttles the
    whole column by the sum of the relative destruction over its 660
    components and ends percent-level wrong; the geometric mean stays at
    the tolerance."""
    dom = cases.ragged_columns(ncol=400, seed=4)
    sub = dom.subset([71])
    sub.params.arith = 1
    assert sub.nlayers[0] == 60
    x0 = sub.conc0.copy()
    ref = x0.copy()
    assert O.Oracle(sub, sub.params.si()).run(ref, 0.01, 60000, RUNGE_KUTTA_4) == 0
    err = {}
    for w in (0, 1):
        q = dataclasses.replace(sub.params, emp_weights=w)
        y = x0.copy()
        o = O.Oracle(sub, q.si())
        assert o.emp_integrate(y, 600.0) == 0
        assert np.all(y >= 0.0)
        err[w] = float(np.max(np.abs(y - ref) / (1e-6 + np.abs(ref))))
    print("max rel error: product", err[0], "geometric", err[1])
    assert err[1] <= 1e-5
    assert err[0] >= 1e-3


@pytest.mark.parametrize("weights", [0, 1])
@pytest.mark.parametrize("method", [EMP1, EMP2])
def test_emp_positive_and_conservative_at_huge_steps(method, weights):
    d = cases.plain_columns(ni=2, nj=2, K=15, seed=12)
    d.params.clamp = 0
    d.params.emp_weights = weights
    p = d.params.si()
    x = d.conc0.copy()
    orc = O.Oracle(d, p)
    inv0 = orc.inventory(x)
    for dt in (3600.0, 86400.0, 30 * 86400.0):
        assert orc.ode_solver(x, dt, method) == 0
        assert np.all(x >= 0.0)
    res = budget_residuals(inv0, orc.inventory(x), orc.budget())
    for e, r in res.items():
        assert r.max() <= 1e-13, (e, r.max())


@pytest.mark.parametrize("weights", [0, 1])
def test_emp_observed_order(weights):
    """Two linear decays in a closed one-layer column (no transport at all):
    EMP1 error ~ dt, EMP2 ~ dt^2."""
    d = _decay_domain(K=1, ncol=1)
    d.conc0[1] = d.conc0[0] * 0.5
    d.params.bcup_dis = 3
    d.params.emp_weights = weights
    p = d.params.si()
    t_end = 40 * 86400.0
    fT = O.oc_exp(-4500.0 * (1.0 / (d.temp[0, 0] + 273.0) - 1.0 / 288.0))
    exact = d.conc0[0] * math.exp(-fT * p["rL"] * t_end)
    orders = {}
    for method in (EMP1, EMP2):
        errs = []
        for n in (20, 40, 80):
            x = d.conc0.copy()
            orc = O.Oracle(d, p)
            for _ in range(n):
                orc.ode_solver(x, t_end / n, method)
            errs.append(np.max(np.abs(x[0] - exact) / exact))
        orders[method] = math.log2(errs[1] / errs[2])
    assert 0.85 < orders[EMP1] < 1.15, orders
    assert 1.8 < orders[EMP2] < 2.2, orders


@pytest.mark.parametrize("weights", [0, 1])
def test_emp_adaptive_budget_and_positivity(weights):
    full = cases.shelf_sea(n=30, days=2, seed=5)
    dom = full.subset(np.flatnonzero(full.mask == 0)[:24])
    dom.params.emp_weights = weights
    p = dom.params.si()
    orc = O.Oracle(dom, p)
    x = dom.conc0.copy()
    spans = cases.record_spans(0.0, 2 * 86400.0, dom.nrec, dom.t0, dom.dt_rec)
    assert orc.integrate_series(x, 0.0, 2 * 86400.0, spans, nthreads=4) == 0
    assert np.all(x >= 0.0)
    assert np.all(orc.diag["accepted"] > 48)
    res = budget_residuals(orc.inventory(dom.conc0), orc.inventory(x), orc.budget())
    for e, r in res.items():
        assert r.max() <= 1e-12, (e, r.max())


# ---------------------------------------------------------------------------
# elementary functions
# ---------------------------------------------------------------------------

def test_exp_within_two_ulp():
    xs = np.concatenate([np.linspace(-700, 700, 2001), np.linspace(-1, 1, 2001), [-0.62, 0.27, 0.0]])
    for x in xs:
        e = O.oc_exp(float(x))
        ref = math.exp(float(x))
        assert abs(e - ref) <= 2 * math.ulp(ref), (x, e, ref)


def test_rsqrt():
    for x in np.geomspace(1e-30, 1e30, 301):
        assert math.isclose(O.oc_rsqrt(float(x)), 1.0 / math.sqrt(float(x)), rel_tol=1e-15)


def test_assemble_forcing_closed_form():
    """get_boundary_conditions (fabm_sediment_component.F90:1865-2030):
    particulate flux -c w, solute bottom water copied, bcup = 1 flux from the
    top layer by the reference formula, absent fields untouched, land
    columns untouched."""
    dom = cases.plain_columns(ni=2, nj=2, K=12, seed=3)
    dom.mask[3] = 1
    p = dom.params.si()
    p["bcup_dis"] = 1
    orc = O.Oracle(dom, p)
    rng = np.random.default_rng(5)
    x = dom.conc0.copy()
    x[:, :] = rng.uniform(0.1, 100.0, size=x.shape)
    t0, c0, f0 = (a.copy() for a in dom.record(0))
    ts = rng.uniform(2.0, 25.0, size=4)
    cs = rng.uniform(0.0, 300.0, size=(11, 4))
    ws = -rng.uniform(1e-6, 1e-4, size=(11, 4))
    present = np.ones(11, np.int32)
    present[7] = 0  # NH4 field absent
    t, c, f = orc.assemble_forcing(x, present, ts, cs, ws)
    off = dom.off
    for col in range(4):
        if col == 3:
            assert t[col] == t0[col] and np.array_equal(c[:, col], c0[:, col]) and np.array_equal(f[:, col], f0[:, col])
            continue
        assert t[col] == ts[col]
        top = off[col]
        for n in range(11):
            if n == 7:
                assert c[n, col] == c0[n, col] and f[n, col] == f0[n, col]
            elif n in (0, 1, 2, 3, 4, 10):
                assert f[n, col] == -(cs[n, col] * ws[n, col])
            else:
                assert c[n, col] == cs[n, col]
                B, D0 = dom.biot[col], p["D0"]
                ref = -(x[n, top] - cs[n, col]) / dom.dz[top] * (B + D0 + ts[col] * 0.035) * dom.phi[top] / 86400.0 / 1e4
                assert f[n, col] == pytest.approx(ref, rel=1e-14, abs=0.0)
    # bcup_dis = 2: solute fluxes reset to zero
    p["bcup_dis"] = 2
    orc = O.Oracle(dom, p)
    _, _, f2 = orc.assemble_forcing(x, present, ts, cs, ws)
    assert np.all(f2[[5, 6, 8, 9]][:, :3] == 0.0)


def test_bcup4_cascade_against_the_reference_loop():
    """diff3d BcUp = 4 (fabm_sediment_driver.F90:791-802): the oracle's
    right-hand side minus the BcUp = 1 one equals the divergence of the
    cascade increments computed here by the reference loop, for columns
    where the cascade stops inside the column and where it passes the
    bottom."""
    from sedcol import grid
    rng = np.random.Generator(np.random.Philox(2))
    nl = np.array([1, 3, 9, 20], dtype=np.int32)
    s = cases.draw_scalars(rng, nl.size)
    dom = cases.build_columns(nl, 0.2, s)
    off = dom.off
    fcap = np.empty(dom.ncell)
    pom = [1e4, 1.0, 30.0, 5e2]
    for c in range(dom.ncol):
        sl = slice(off[c], off[c + 1])
        fcap[sl] = grid.flux_cap(dom.phi[sl], dom.dz[sl], pom[c])
    x = dom.conc0.copy()
    x[0:5] = rng.uniform(10.0, 1e3, size=x[0:5].shape)
    p4 = dom.params.si()
    p4["bcup_part"] = 4
    r4, f4 = O.Oracle(dom, p4, fcap=fcap).rhs(x)
    r1, f1 = O.Oracle(dom, dict(p4, bcup_part=1), fcap=fcap).rhs(x)
    fup = dom.record(0)[2]
    depths = []
    for c in range(dom.ncol):
        K = int(nl[c])
        cap = fcap[off[c]:off[c + 1]]
        for n in (0, 1, 2, 3, 4, 10):
            inc = np.zeros(K + 1)  # added to interface k (0 = top)
            k = 1
            rest = fup[n, c] - cap[0]
            while rest > 0 and k < K:
                inc[k] += rest
                rest -= cap[k]
                k += 1
            if k >= K:
                inc[K - 1] += rest
            if n == 0:
                depths.append(k)
            phi = dom.phi[off[c]:off[c + 1]]
            dz = dom.dz[off[c]:off[c + 1]]
            expect = (inc[:K] - inc[1:K + 1]) / ((1.0 - phi) * dz) * (1.0 - phi) / phi
            got = r4[n, off[c]:off[c + 1]] - r1[n, off[c]:off[c + 1]]
            scale = np.abs(r1[n, off[c]:off[c + 1]]).max() + np.abs(expect).max()
            assert np.allclose(got, expect, rtol=1e-9, atol=1e-12 * scale), (c, n)
    assert min(depths) < max(nl) and any(d >= K for d, K in zip(depths, nl))
    # the top flux is the prescribed one, except for a single-layer column,
    # where the overflow lands on the top interface itself (k >= K at once)
    assert np.array_equal(f4[:, 1:], f1[:, 1:])
    assert np.allclose(f4[[0, 1, 2, 3, 4, 10], 0], fup[[0, 1, 2, 3, 4, 10], 0] + (fup[[0, 1, 2, 3, 4, 10], 0] - fcap[0]),
                       rtol=1e-15)


def test_log_within_two_ulp():
    xs = np.concatenate([np.geomspace(1e-320, 1e300, 4001), np.linspace(0.5, 2.0, 2001),
                         [1.0, 2.0, 0.7071067811865476, 1.4142135623730951, 5e-324, 1.7976931348623157e308]])
    for x in xs:
        v = O.oc_log(float(x))
        ref = math.log(float(x))
        assert abs(v - ref) <= 2 * math.ulp(ref) + (1e-300 if ref == 0.0 else 0.0), (x, v, ref)
    assert O.oc_log(0.0) == -math.inf and math.isnan(O.oc_log(-1.0)) and O.oc_log(math.inf) == math.inf


ZW = (0.11, 0.2, 0.6, 0.22, 1.334, 1000.0)  # k_l, L1, L2, beta, b, dry density (fabm_sediment_driver.F90:226-231)


def _zw_bfac_numpy(dom, x, zw):
    """Bioturbation factor of profile 3 (fabm_sediment_driver.F90:618-645),
    restated with numpy's exp and log."""
    k_l, L1, L2, beta, b, rho = zw
    fl, fs = 1.0 / 1.2 * 12.01 / rho / 1000.0, 1.0 / 6.0 * 12.01 / rho / 1000.0
    out = np.empty(dom.ncell)
    off = dom.off
    for c in range(dom.ncol):
        sl = slice(off[c], off[c + 1])
        dz, phi = dom.dz[sl], dom.phi[sl]
        zc = np.cumsum(dz) - dz / 2.0
        cum = dz[:-1].sum()  # the light loop's last cumdepth: layers 1..K-1
        wt = fl * phi / (1.0 - phi) * x[0, sl] + fs * phi / (1.0 - phi) * x[1, sl]
        avg = np.sum(dz * wt) / cum
        bio = wt * np.exp(zc * 100.0 * k_l) * avg / (L1 + L2 * np.exp(zc * 200.0 * k_l))
        out[sl] = beta * bio ** b / wt
    return out


def test_profile3_bioturbation_against_numpy():
    """With profile 3 the oracle's right-hand side equals the constant-profile
    one whose factor is the numpy restatement, at 15 degC (f_T = 1) and
    B = 1 (profile 3 sets both to 1)."""
    dom = cases.plain_columns(ni=2, nj=2, K=20, seed=7)
    dom.temp[:] = 15.0
    dom.biot[:] = 1.0
    p = dom.params.si()
    rng = np.random.default_rng(3)
    x = dom.conc0.copy()
    x[0:5] = rng.uniform(1.0, 5e3, size=x[0:5].shape)
    r3, f3 = O.Oracle(dom, p, zw=ZW).rhs(x)
    d2 = dataclasses.replace(dom, bfac=_zw_bfac_numpy(dom, x, ZW))
    r0, f0 = O.Oracle(d2, p).rhs(x)
    assert np.allclose(r3, r0, rtol=1e-11, atol=1e-300)
    assert np.allclose(f3, f0, rtol=1e-11, atol=1e-300)
    # the factor matters: the set_geometry profile gives another result
    r1, _ = O.Oracle(dom, p).rhs(x)
    assert not np.allclose(r1, r3, rtol=1e-6)
