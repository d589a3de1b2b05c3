"""GPU parity: libsedcol.so (through the C ABI) against the CPU oracle.

Tolerances: 1e-10 relative per RK4 step, 1e-8 after a
simulated year, C/N/P budget closure 1e-12 relative, non-negative state
under EMP.  The kernels and the oracle execute the same IEEE operation
sequence, so most comparisons here are bitwise; the stated tolerance is
asserted and bitwise equality is asserted where the design guarantees it.
"""
import numpy as np
import pytest

from sedcol import cases, capi
from sedcol.params import EULER, RUNGE_KUTTA_4, RUNGE_KUTTA_4_38, ADAPTIVE_EULER, EMP1, EMP2
import oracle as O
from helpers import max_rel, budget_residuals

pytestmark = pytest.mark.gpu


def _perturbed(dom, seed=3, lo=0.5, hi=1.5):
    rng = np.random.default_rng(seed)
    x = dom.conc0.copy()
    # give particulates and ODU non-trivial profiles
    x[0:5] = rng.uniform(0.0, 1.0e4, size=x[0:5].shape)
    x[8] = rng.uniform(0.0, 50.0, size=x[8].shape)
    x[10] = rng.uniform(0.0, 10.0, size=x[10].shape)
    return x * rng.uniform(lo, hi, size=x.shape)


ARITHS = [pytest.param(0, id="reference"), pytest.param(1, id="fast")]


@pytest.mark.parametrize("arith", ARITHS)
def test_rhs_bitwise(gpu_ready, arith):
    dom = cases.plain_columns(seed=0)
    dom.params.arith = arith
    p = dom.params.si()
    x = _perturbed(dom)
    with capi.SedimentColumns.from_domain(dom) as g:
        rg, fg = g.get_rhs(x)
    ro, fo = O.Oracle(dom, p).rhs(x)
    print("rhs max rel", max_rel(rg, ro), "flux max rel", max_rel(fg, fo), flush=True)
    assert np.array_equal(rg, ro)
    assert np.array_equal(fg, fo)


@pytest.mark.parametrize("bcup_dis,bcup_part,bcdown", [(1, 1, 1), (3, 3, 2), (2, 2, 3)])
def test_rhs_boundary_kinds(gpu_ready, bcup_dis, bcup_part, bcdown):
    dom = cases.plain_columns(ni=2, nj=3, K=17, seed=5)
    dom.params.bcup_dis, dom.params.bcup_part, dom.params.bcdown = bcup_dis, bcup_part, bcdown
    p = dom.params.si()
    x = _perturbed(dom, seed=9)
    with capi.SedimentColumns.from_domain(dom) as g:
        rg, fg = g.get_rhs(x)
    ro, fo = O.Oracle(dom, p).rhs(x)
    assert np.array_equal(rg, ro)
    assert np.array_equal(fg, fo)


@pytest.mark.parametrize("arith", ARITHS)
@pytest.mark.parametrize("method", [EULER, RUNGE_KUTTA_4, RUNGE_KUTTA_4_38, EMP1, EMP2])
def test_one_step(gpu_ready, method, arith):
    dom = cases.plain_columns(seed=1)
    dom.params.arith = arith
    p = dom.params.si()
    x0 = _perturbed(dom, seed=4)
    xg = x0.copy()
    with capi.SedimentColumns.from_domain(dom) as g:
        fg = g.ode_solver(xg, 360.0, method)
    xo = x0.copy()
    orc = O.Oracle(dom, p)
    assert orc.ode_solver(xo, 360.0, method) == 0
    err = max_rel(xg, xo)
    print(f"method {method}: one-step max rel {err:.3e}", flush=True)
    assert err <= 1e-10
    assert np.array_equal(xg, xo)
    assert np.array_equal(fg, orc.flux)


@pytest.mark.parametrize("arith", ARITHS)
def test_rk4_30_days(gpu_ready, arith):
    """Plain 4x4x30 RK4 case, dt = 360 s, 7200 steps in one persistent launch."""
    dom = cases.plain_columns(seed=0)
    dom.params.arith = arith
    p = dom.params.si()
    with capi.SedimentColumns.from_domain(dom) as g:
        g.run(360.0, 7200, RUNGE_KUTTA_4)
        xg = g.get_state()
        bg = g.budget()
        invg = g.inventory()
        dg = g.diagnostics()
    orc = O.Oracle(dom, p)
    xo = dom.conc0.copy()
    assert orc.run(xo, 360.0, 7200, RUNGE_KUTTA_4, nthreads=8) == 0
    err = max_rel(xg, xo)
    print(f"RK4 30 d max rel {err:.3e}", flush=True)
    assert err <= 1e-8
    assert np.array_equal(xg, xo)
    assert np.allclose(bg, orc.budget(), rtol=1e-12, atol=0)
    assert np.array_equal(invg, orc.inventory(xo))
    assert np.all(dg["accepted"] == 7200)
    res = budget_residuals(orc.inventory(dom.conc0), invg, bg)
    for e, r in res.items():
        print(f"budget {e}: max rel closure {r.max():.3e}", flush=True)
        assert r.max() <= 1e-12


@pytest.mark.parametrize("arith", ARITHS)
@pytest.mark.parametrize("method", [EMP1, EMP2])
def test_geometric_weights_one_step(gpu_ready, method, arith):
    """Fixed EMP steps with geometric-mean Patankar weights, bitwise."""
    dom = cases.plain_columns(seed=1)
    dom.params.arith = arith
    dom.params.emp_weights = 1
    p = dom.params.si()
    x0 = _perturbed(dom, seed=4)
    for dt in (360.0, 86400.0):
        xg = x0.copy()
        with capi.SedimentColumns.from_domain(dom) as g:
            fg = g.ode_solver(xg, dt, method)
            dg = g.diagnostics()
        xo = x0.copy()
        orc = O.Oracle(dom, p)
        assert orc.ode_solver(xo, dt, method) == 0
        assert np.array_equal(xg, xo)
        assert np.array_equal(fg, orc.flux)
        assert np.array_equal(dg["newton"], orc.diag["newton"])
        assert np.all(xg >= 0.0)


def test_rk4_windows_equal_single_launch(gpu_ready):
    dom = cases.plain_columns(ni=2, nj=2, seed=2)
    with capi.SedimentColumns.from_domain(dom) as g:
        g.run(360.0, 500, RUNGE_KUTTA_4, steps_per_launch=0)
        a = g.get_state()
    with capi.SedimentColumns.from_domain(dom) as g:
        g.run(360.0, 500, RUNGE_KUTTA_4, steps_per_launch=7)
        b = g.get_state()
        ms, launches = g.last_timing()
    assert launches == (500 + 6) // 7
    assert np.array_equal(a, b)


@pytest.mark.parametrize("nij", [pytest.param(2, id="one_workgroup"), pytest.param(6, id="multi_launch")])
def test_adaptive_euler_global_rule(gpu_ready, nij):
    """ADAPTIVE_EULER quarters the step for every column when any cell
    violates the relative-change test (solver_library.F90:119-129).  4
    columns run in one workgroup (the domain-wide vote inside the kernel),
    36 (more than one workgroup of 16-lane groups holds) in one launch per
    stage with the vote read by the host; both bitwise like the oracle."""
    dom = cases.plain_columns(ni=nij, nj=nij, K=12, seed=7)
    dom.params.dt_min = 1.0
    p = dom.params.si()
    x0 = dom.conc0.copy()
    x0[0:2] = 5.0e4  # labile detritus present from the start: fast O2 drawdown
    xg = x0.copy()
    with capi.SedimentColumns.from_domain(dom) as g:
        g.set_state(xg)
        g.run(3600.0, 3, ADAPTIVE_EULER)
        xg = g.get_state()
        lmg = g.last_min_dt()
        acc_g = g.diagnostics()["accepted"]
    orc = O.Oracle(dom, p)
    xo = x0.copy()
    for _ in range(3):
        orc.ode_solver(xo, 3600.0, ADAPTIVE_EULER)
    assert np.array_equal(xg, xo)
    assert lmg == orc.last_min_dt[0]
    assert lmg < 3600.0  # the test case does subcycle
    assert np.all(acc_g == orc.diag["accepted"])


@pytest.mark.parametrize("arith", ARITHS)
def test_emp_adaptive_shelf_subset(gpu_ready, arith):
    """Adaptive EMP across daily forcing records on shelf-sea columns."""
    full = cases.shelf_sea(n=40, days=3, seed=11)
    wet = np.flatnonzero(full.mask == 0)[:64]
    dom = full.subset(wet)
    dom.params.arith = arith
    p = dom.params.si()
    t1 = 3 * 86400.0
    with capi.SedimentColumns.from_domain(dom) as g:
        g.integrate(0.0, t1)
        xg = g.get_state()
        bg = g.budget()
        dg = g.diagnostics()
        fg = g.fluxes()
    orc = O.Oracle(dom, p)
    xo = dom.conc0.copy()
    spans = cases.record_spans(0.0, t1, dom.nrec, dom.t0, dom.dt_rec)
    assert orc.integrate_series(xo, 0.0, t1, spans, nthreads=8) == 0
    print("EMP accepted", dg["accepted"].sum(), "rejected", dg["rejected"].sum(),
          "max rel", max_rel(xg, xo), flush=True)
    assert np.array_equal(xg, xo)
    assert np.array_equal(dg["accepted"], orc.diag["accepted"])
    assert np.array_equal(dg["rejected"], orc.diag["rejected"])
    assert np.array_equal(fg, orc.flux)
    assert np.allclose(bg, orc.budget(), rtol=1e-12, atol=1e-300)
    assert np.all(xg >= 0.0)
    res = budget_residuals(orc.inventory(dom.conc0), orc.inventory(xg), bg)
    for e, r in res.items():
        assert r.max() <= 1e-12, (e, r.max())


@pytest.mark.parametrize("weights", [pytest.param(0, id="product"), pytest.param(1, id="geometric")])
@pytest.mark.parametrize("arith", ARITHS)
def test_emp_ragged(gpu_ready, arith, weights):
    dom = cases.ragged_columns(ncol=96, seed=4)
    assert set(np.unique(np.where(dom.nlayers <= 16, 16, np.where(dom.nlayers <= 32, 32, 64)))) == {16, 32, 64}
    dom.params.arith = arith
    dom.params.emp_weights = weights
    p = dom.params.si()
    span = 6 * 3600.0
    with capi.SedimentColumns.from_domain(dom) as g:
        g.integrate(0.0, span)
        xg = g.get_state()
    orc = O.Oracle(dom, p)
    xo = dom.conc0.copy()
    assert orc.emp_integrate(xo, span, nthreads=8) == 0
    assert np.array_equal(xg, xo)
    assert np.all(xg >= 0.0)


def _edge_columns():
    """One column at each layer count where the lane mapping changes: a
    single layer, two, the 16/32/64-lane bin edges and the maximum of 64."""
    K = np.array([1, 2, 15, 16, 17, 31, 32, 33, 63, 64], dtype=np.int32)
    rng = np.random.Generator(np.random.Philox(11))
    s = cases.draw_scalars(rng, K.size)
    from sedcol.params import SedParams
    p = SedParams(bcup_dis=2, bcup_part=1, bcdown=3, clamp=0, emp_rtol=1e-6, emp_atol=1e-6,
                  emp_dt_min=1.0, emp_dt_max=3600.0)
    return cases.build_columns(K, np.linspace(0.05, 0.5, K.size), s, params=p)


@pytest.mark.parametrize("weights", [pytest.param(0, id="product"), pytest.param(1, id="geometric")])
@pytest.mark.parametrize("arith", ARITHS)
def test_layer_count_edges(gpu_ready, arith, weights):
    dom = _edge_columns()
    dom.params.arith = arith
    dom.params.emp_weights = weights
    p = dom.params.si()
    # RK4, one simulated day in one launch
    with capi.SedimentColumns.from_domain(dom) as g:
        g.run(360.0, 240, RUNGE_KUTTA_4)
        xg = g.get_state()
        bg = g.budget()
    orc = O.Oracle(dom, p)
    xo = dom.conc0.copy()
    assert orc.run(xo, 360.0, 240, RUNGE_KUTTA_4, nthreads=8) == 0
    assert np.array_equal(xg, xo)
    assert np.allclose(bg, orc.budget(), rtol=1e-12, atol=1e-300)
    # adaptive EMP, six hours
    with capi.SedimentColumns.from_domain(dom) as g:
        g.integrate(0.0, 6 * 3600.0)
        xg = g.get_state()
        dg = g.diagnostics()
    orc = O.Oracle(dom, p)
    xo = dom.conc0.copy()
    assert orc.emp_integrate(xo, 6 * 3600.0, nthreads=8) == 0
    print("edge columns: EMP attempts", (dg["accepted"] + dg["rejected"]).tolist(), flush=True)
    assert np.array_equal(xg, xo)
    assert np.array_equal(dg["accepted"], orc.diag["accepted"])
    assert np.all(xg >= 0.0)


def test_fortran_layout_matches_packed(gpu_ready):
    dom = cases.plain_columns(ni=3, nj=2, K=20, seed=6)
    p = dom.params.si()
    K, ncol = 20, dom.ncol
    cellf = lambda a: np.ascontiguousarray(a.reshape(ncol, K).T)
    with capi.SedimentColumns(np.full(ncol, K), layout="fortran") as g:
        g.set_params(p)
        g.set_geometry(cellf(dom.dz), cellf(dom.dzc), cellf(dom.phi), cellf(dom.phiI), cellf(dom.bfac), dom.biot)
        g.set_forcing(*dom.record(0))
        xf = dom.fortran_state(dom.conc0)
        g.set_state(xf)
        g.run(360.0, 100, RUNGE_KUTTA_4)
        af = g.get_state()
    with capi.SedimentColumns.from_domain(dom) as g:
        g.run(360.0, 100, RUNGE_KUTTA_4)
        ap = g.get_state()
    assert np.array_equal(dom.packed_state(af), ap)


def test_masked_columns_untouched(gpu_ready):
    full = cases.shelf_sea(n=12, days=2, seed=3)
    assert full.mask.any() and not full.mask.all()
    with capi.SedimentColumns.from_domain(full) as g:
        g.integrate(0.0, 86400.0)
        x = g.get_state()
    with capi.SedimentColumns.from_domain(full, series=False) as g:
        r, f = g.get_rhs(full.conc0)
    land = np.repeat(full.mask, full.nlayers) != 0
    assert np.all(x[:, land] == 1.0e20)
    assert np.all(r[:, land] == 0.0)
    assert np.all(x[:, ~land] >= 0.0)


def test_errors_are_status_codes(gpu_ready):
    dom = cases.plain_columns(ni=1, nj=2, K=8, seed=0)
    g = capi.SedimentColumns(dom.nlayers)
    bad = dom.phi.copy()
    bad[3] = 1.2
    with pytest.raises(capi.SedcolError) as e:
        g.set_geometry(dom.dz, dom.dzc, bad, dom.phiI, dom.bfac, dom.biot)
    assert e.value.code == capi.SEDCOL_EDOMAIN
    p = dom.params.si()
    p["bcup_part"] = 5
    with pytest.raises(capi.SedcolError) as e:
        g.set_params(p)
    assert e.value.code == capi.SEDCOL_EINVAL
    with pytest.raises(capi.SedcolError) as e:
        g.run(360.0, 1, RUNGE_KUTTA_4)
    assert e.value.code == capi.SEDCOL_EINVAL
    g.close()


def test_nan_state_reported(gpu_ready):
    dom = cases.plain_columns(ni=1, nj=2, K=8, seed=0)
    x = dom.conc0.copy()
    x[5, 3] = np.nan
    with capi.SedimentColumns.from_domain(dom) as g:
        g.set_state(x)
        with pytest.raises(capi.SedcolError) as e:
            g.run(360.0, 10, RUNGE_KUTTA_4)
        assert e.value.code == capi.SEDCOL_ENAN
        assert g.diagnostics()["status"][0] == 1


@pytest.mark.parametrize("bcup_dis", [1, 2])
def test_assemble_forcing_bitwise(gpu_ready, bcup_dis):
    """Boundary assembly from the fields at the soil surface on the device
    against the oracle: top fluxes returned, then the right-hand side under
    the assembled forcing (temperature and bottom water)."""
    dom = cases.ragged_columns(ncol=40, seed=8)
    dom.mask[5] = 1
    dom.params.bcup_dis = bcup_dis
    p = dom.params.si()
    rng = np.random.default_rng(9)
    x = _perturbed(dom, seed=10)
    ts = rng.uniform(2.0, 25.0, size=dom.ncol)
    cs = rng.uniform(0.0, 300.0, size=(11, dom.ncol))
    ws = -rng.uniform(1e-6, 1e-4, size=(11, dom.ncol))
    present = np.ones(11, np.int32)
    present[9] = 0
    orc = O.Oracle(dom, p)
    _, _, fo = orc.assemble_forcing(x, present, ts, cs, ws)
    with capi.SedimentColumns.from_domain(dom) as g:
        g.set_state(x)
        fg = g.assemble_forcing(present, ts, cs, ws)
        rg, flg = g.get_rhs(x)
    wet = dom.mask == 0
    assert np.array_equal(fg[:, wet], fo[:, wet])
    ro, flo = orc.rhs(x)
    assert np.array_equal(rg, ro)
    assert np.array_equal(flg, flo)


def _bcup4_domain(seed=12):
    """Columns of 1..64 layers with flux caps from tiny (the whole top flux
    cascades past the bottom) to large (no cascade)."""
    from sedcol import grid
    rng = np.random.Generator(np.random.Philox(seed))
    nl = np.array([1, 2, 3, 5, 8, 16, 17, 24, 30, 33, 48, 64] * 2, dtype=np.int32)
    s = cases.draw_scalars(rng, nl.size)
    dom = cases.build_columns(nl, 0.2, s)
    dom.params.bcup_part = 4
    off = dom.off
    fcap = np.empty(dom.ncell)
    for c in range(dom.ncol):
        sl = slice(off[c], off[c + 1])
        pom = 10.0 ** rng.uniform(-1.0, 4.5)  # mmol m-3 d-1 scale of pom_flux_max
        fcap[sl] = grid.flux_cap(dom.phi[sl], dom.dz[sl], pom)
    return dom, fcap


@pytest.mark.parametrize("arith", ARITHS)
def test_bcup4_rhs_and_emp_bitwise(gpu_ready, arith):
    """diff3d BcUp = 4 (capped top flux cascading downward,
    fabm_sediment_driver.F90:791-802) on the GPU against the oracle: one
    right-hand side, then 6 hours of adaptive EMP (the fast form falls back
    to the reference form for BcUp = 4)."""
    dom, fcap = _bcup4_domain()
    dom.params.arith = arith
    p = dom.params.si()
    x = _perturbed(dom, seed=13)
    orc = O.Oracle(dom, p, fcap=fcap)
    ro, fo = orc.rhs(x)
    with capi.SedimentColumns.from_domain(dom) as g:
        g.set_flux_cap(fcap)
        rg, fg = g.get_rhs(x)
        assert np.array_equal(rg, ro)
        assert np.array_equal(fg, fo)
        g.integrate(0.0, 6 * 3600.0)
        xg = g.get_state()
    xo = dom.conc0.copy()
    assert orc.emp_integrate(xo, 6 * 3600.0, nthreads=8) == 0
    assert np.array_equal(xg, xo)
    assert np.all(xg >= 0.0)
    # the cascade is active: with the cap removed the right-hand side differs
    ro1, _ = O.Oracle(dom, dict(p, bcup_part=1)).rhs(x)
    assert not np.array_equal(ro1, ro)


def test_extensions_with_geometric_weights(gpu_ready):
    """The kernel instances of the transport extensions (BcUp = 4, profile 3)
    with geometric-mean Patankar weights: 6 hours of adaptive EMP each,
    bitwise against the oracle."""
    dom, fcap = _bcup4_domain()
    dom.params.emp_weights = 1
    p = dom.params.si()
    with capi.SedimentColumns.from_domain(dom) as g:
        g.set_flux_cap(fcap)
        g.integrate(0.0, 6 * 3600.0)
        xg = g.get_state()
    xo = dom.conc0.copy()
    assert O.Oracle(dom, p, fcap=fcap).emp_integrate(xo, 6 * 3600.0, nthreads=8) == 0
    assert np.array_equal(xg, xo)
    rng = np.random.Generator(np.random.Philox(14))
    nl = np.array([2, 17, 33, 64] * 2, dtype=np.int32)
    dom = cases.build_columns(nl, 0.2, cases.draw_scalars(rng, nl.size))
    dom.params.emp_weights = 1
    p = dom.params.si()
    x = _perturbed(dom, seed=15)
    with capi.SedimentColumns.from_domain(dom) as g:
        g.set_bioturbation_zw()
        g.set_state(x)
        g.integrate(0.0, 6 * 3600.0)
        xg = g.get_state()
    xo = x.copy()
    assert O.Oracle(dom, p, zw=capi.ZW_DEFAULTS).emp_integrate(xo, 6 * 3600.0, nthreads=8) == 0
    assert np.array_equal(xg, xo)
    assert np.all(xg >= 0.0)


def test_bcup4_needs_flux_cap(gpu_ready):
    dom, fcap = _bcup4_domain()
    with capi.SedimentColumns.from_domain(dom) as g:
        with pytest.raises(capi.SedcolError) as e:
            g.integrate(0.0, 3600.0)
        assert e.value.code == capi.SEDCOL_EINVAL
        bad = fcap.copy()
        bad[4] = -1.0
        with pytest.raises(capi.SedcolError) as e:
            g.set_flux_cap(bad)
        assert e.value.code == capi.SEDCOL_EDOMAIN


@pytest.mark.parametrize("arith", ARITHS)
def test_profile3_bioturbation_bitwise(gpu_ready, arith):
    """Bioturbation profile 3 (Zhang & Wirtz, fabm_sediment_driver.F90:
    618-645): the factor recomputed from the TOC profile on every right-hand
    side, on the GPU against the oracle: one right-hand side, then 6 hours of
    adaptive EMP from a state with detritus (the fast form falls back to the
    reference form)."""
    from sedcol import grid
    rng = np.random.Generator(np.random.Philox(14))
    nl = np.array([2, 3, 9, 16, 17, 30, 33, 64] * 3, dtype=np.int32)
    s = cases.draw_scalars(rng, nl.size)
    dom = cases.build_columns(nl, 0.2, s)
    dom.params.arith = arith
    p = dom.params.si()
    x = _perturbed(dom, seed=15)
    orc = O.Oracle(dom, p, zw=capi.ZW_DEFAULTS)
    ro, fo = orc.rhs(x)
    with capi.SedimentColumns.from_domain(dom) as g:
        g.set_bioturbation_zw()
        rg, fg = g.get_rhs(x)
        assert np.array_equal(rg, ro)
        assert np.array_equal(fg, fo)
        g.set_state(x)
        g.integrate(0.0, 6 * 3600.0)
        xg = g.get_state()
        g.set_bioturbation_zw(None)
        r_off, _ = g.get_rhs(x)
    xo = x.copy()
    assert orc.emp_integrate(xo, 6 * 3600.0, nthreads=8) == 0
    assert np.array_equal(xg, xo)
    # switched off: the set_geometry factor again
    assert np.array_equal(r_off, O.Oracle(dom, p).rhs(x)[0])
    assert not np.array_equal(r_off, ro)


@pytest.mark.parametrize("arith", ARITHS)
def test_rk4_one_year(gpu_ready, arith):
    """The plain 4x4x30 case run for one simulated year (87 600 RK4 steps of
    360 s, minimum clamp after each): relative 1e-8 against the oracle
    (bitwise in fact) and C/N/P budget closure to 1e-12."""
    dom = cases.plain_columns(seed=0)
    dom.params.arith = arith
    p = dom.params.si()
    nsteps = 365 * 240
    with capi.SedimentColumns.from_domain(dom) as g:
        g.run(360.0, nsteps, RUNGE_KUTTA_4)
        xg = g.get_state()
        bg = g.budget()
        invg = g.inventory()
    orc = O.Oracle(dom, p)
    xo = dom.conc0.copy()
    assert orc.run(xo, 360.0, nsteps, RUNGE_KUTTA_4, nthreads=8) == 0
    err = max_rel(xg, xo)
    print(f"RK4 1 year max rel {err:.3e}", flush=True)
    assert err <= 1e-8
    assert np.array_equal(xg, xo)
    res = budget_residuals(orc.inventory(dom.conc0), invg, bg)
    for e, r in res.items():
        print(f"budget {e}: max rel closure {r.max():.3e}", flush=True)
        assert r.max() <= 1e-12


@pytest.mark.parametrize("weights", [pytest.param(0, id="product"), pytest.param(1, id="geometric")])
def test_shelf_sea_full_size_properties(gpu_ready, weights):
    """The full 300 x 300 shelf-sea domain (53 211 wet columns), first
    simulated day from the cold state under adaptive EMP: every
    concentration finite and non-negative, land untouched, and the C/N/P
    budget of every column closed to 1e-12 relative (inventory change =
    integrated top fluxes - DIC/N2 release + clamp mass)."""
    dom = cases.shelf_sea(n=300, days=2, seed=0)
    dom.params.arith = 1
    dom.params.emp_weights = weights
    land = np.repeat(dom.mask, dom.nlayers) != 0
    with capi.SedimentColumns.from_domain(dom) as g:
        inv0 = g.inventory()
        g.integrate(0.0, 86400.0)
        x = g.get_state()
        b = g.budget()
        inv1 = g.inventory()
        d = g.diagnostics()
    wet = dom.mask == 0
    assert np.all(np.isfinite(x[:, ~land])) and np.all(x[:, ~land] >= 0.0)
    assert np.array_equal(x[:, land], dom.conc0[:, land])
    assert np.all(d["status"][wet] == 0) and np.all(d["accepted"][wet] > 0)
    res = budget_residuals(inv0, inv1, b)
    for e, r in res.items():
        print(f"budget {e}: max rel closure {r[wet].max():.3e}", flush=True)
        assert r[wet].max() <= 1e-12


@pytest.mark.parametrize("weights", [pytest.param(0, id="product"), pytest.param(1, id="geometric")])
def test_ragged_full_size_properties(gpu_ready, weights):
    """The full ragged batch (1e5 columns, K U{10..60}, 3.5 M cells), first
    simulated hour from the cold state under adaptive EMP: every
    concentration finite and non-negative, every column accepted steps, and
    the C/N/P budget of every column closed to 1e-12 relative."""
    dom = cases.ragged_columns(ncol=100_000, seed=0)
    dom.params.arith = 1
    dom.params.emp_weights = weights
    with capi.SedimentColumns.from_domain(dom) as g:
        inv0 = g.inventory()
        g.integrate(0.0, 3600.0)
        x = g.get_state()
        b = g.budget()
        inv1 = g.inventory()
        d = g.diagnostics()
    assert np.all(np.isfinite(x)) and np.all(x >= 0.0)
    assert np.all(d["status"] == 0) and np.all(d["accepted"] > 0)
    res = budget_residuals(inv0, inv1, b)
    for e, r in res.items():
        print(f"ragged budget {e}: max rel closure {r.max():.3e}", flush=True)
        assert r.max() <= 1e-12
