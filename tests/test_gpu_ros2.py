"""GPU parity of the linearly implicit stepper (method 7, k_ros_adaptive)
against the oracle's col_ros2_adaptive: states, top fluxes, budgets and
step counts bitwise equal (same IEEE operation sequence), in both
arithmetic forms, across forcing records, at the lane-mapping edges of the
layer count, and through sedcol_step; full-size properties (positivity,
C/N/P closure 1e-12) on the shelf-sea domain."""
import numpy as np
import pytest

from sedcol import cases, capi, spinup
from sedcol.params import ROS2_ADAPTIVE
import oracle as O
from helpers import max_rel, budget_residuals

pytestmark = pytest.mark.gpu

ARITHS = [pytest.param(0, id="reference"), pytest.param(1, id="fast")]


def _compare(dom, xg, dg, bg, fg, orc, xo):
    print("ROS2 accepted", dg["accepted"].sum(), "rejected", dg["rejected"].sum(), "max rel", max_rel(xg, xo),
          flush=True)
    assert np.array_equal(xg, xo)
    assert np.array_equal(dg["accepted"], orc.diag["accepted"])
    assert np.array_equal(dg["rejected"], orc.diag["rejected"])
    assert np.array_equal(fg, orc.flux)
    assert np.array_equal(bg, orc.budget())


@pytest.mark.parametrize("arith", ARITHS)
def test_ros2_shelf_subset_bitwise(gpu_ready, arith):
    full = cases.shelf_sea(n=40, days=3, seed=11)
    dom = full.subset(np.flatnonzero(full.mask == 0)[:64])
    dom.params.arith = arith
    dom.params.adaptive_method = ROS2_ADAPTIVE
    p = dom.params.si()
    t1 = 3 * 86400.0
    with capi.SedimentColumns.from_domain(dom) as g:
        g.integrate(0.0, t1)
        xg, bg, dg, fg = g.get_state(), g.budget(), g.diagnostics(), g.fluxes()
    orc = O.Oracle(dom, p)
    xo = dom.conc0.copy()
    spans = cases.record_spans(0.0, t1, dom.nrec, dom.t0, dom.dt_rec)
    assert orc.integrate_series(xo, 0.0, t1, spans, nthreads=8) == 0
    _compare(dom, xg, dg, bg, fg, orc, xo)
    assert np.all(xg >= 0.0)
    res = budget_residuals(orc.inventory(dom.conc0), orc.inventory(xg), bg)
    for e, r in res.items():
        assert r.max() <= 1e-12, (e, r.max())


@pytest.mark.parametrize("arith", ARITHS)
def test_ros2_layer_count_edges_bitwise(gpu_ready, arith):
    """Spun-up columns (stiff regime) with K at the lane-mapping edges."""
    rng = np.random.Generator(np.random.Philox(5))
    Ks = np.array([1, 2, 5, 11, 12, 15, 16, 17, 31, 32, 33, 63, 64, 30, 30, 40])
    s = spinup.draw_inputs(rng, Ks.size)
    dom = spinup.build(s, Ks)
    dom.params.arith = arith
    p0 = dict(dom.params.si(), clamp=0)
    x0 = dom.conc0.copy()
    O.Oracle(dom, p0).emp_integrate(x0, 5 * 86400.0, nthreads=8)
    dom.conc0 = x0
    dom.params.adaptive_method = ROS2_ADAPTIVE
    p = dom.params.si()
    with capi.SedimentColumns.from_domain(dom) as g:
        g.integrate(0.0, 86400.0)
        xg, bg, dg, fg = g.get_state(), g.budget(), g.diagnostics(), g.fluxes()
    orc = O.Oracle(dom, p)
    xo = x0.copy()
    assert orc.emp_integrate(xo, 86400.0, nthreads=8) == 0
    _compare(dom, xg, dg, bg, fg, orc, xo)


def test_ros2_through_sedcol_step(gpu_ready):
    """sedcol_step with method 7 on a handle whose adaptive stepper is EMP:
    the call uses ROS2 and the handle keeps EMP afterwards."""
    dom = cases.plain_columns(ni=2, nj=2, K=30, seed=4)
    dom.params.clamp = 0
    p = dom.params.si()
    x = dom.conc0.copy()
    with capi.SedimentColumns.from_domain(dom) as g:
        g.ode_solver(x, 6 * 3600.0, ROS2_ADAPTIVE)
        dg = g.diagnostics()
        x_emp = dom.conc0.copy()
        g.reset_diagnostics()
        g.ode_solver(x_emp, 6 * 3600.0, 6)
        de = g.diagnostics()
    xo = dom.conc0.copy()
    orc = O.Oracle(dom, p)
    assert orc.ode_solver(xo, 6 * 3600.0, ROS2_ADAPTIVE) == 0
    assert np.array_equal(x, xo)
    assert np.array_equal(dg["accepted"], orc.diag["accepted"])
    xe = dom.conc0.copy()
    oe = O.Oracle(dom, p)
    assert oe.ode_solver(xe, 6 * 3600.0, 6) == 0
    assert np.array_equal(x_emp, xe)
    assert np.array_equal(de["accepted"], oe.diag["accepted"])


def test_ros2_full_size_properties(gpu_ready):
    """The full 300 x 300 shelf-sea domain, first simulated day from the
    cold state: every concentration non-negative and finite, land
    untouched, C/N/P closure <= 1e-12 for every column."""
    dom = cases.shelf_sea(n=300, days=2, seed=0)
    dom.params.arith = 1
    dom.params.adaptive_method = ROS2_ADAPTIVE
    land = np.repeat(dom.mask, dom.nlayers) != 0
    with capi.SedimentColumns.from_domain(dom) as g:
        inv0 = g.inventory()
        g.integrate(0.0, 86400.0)
        x = g.get_state()
        b = g.budget()
        inv1 = g.inventory()
        d = g.diagnostics()
    wet = dom.mask == 0
    print("ROS2 full size: accepted", d["accepted"][wet].sum(), "rejected", d["rejected"][wet].sum(), flush=True)
    assert np.all(np.isfinite(x[:, ~land])) and np.all(x[:, ~land] >= 0.0)
    assert np.array_equal(x[:, land], dom.conc0[:, land])
    assert np.all(d["status"][wet] == 0)
    res = budget_residuals(inv0, inv1, b)
    for e, r in res.items():
        print(f"budget {e}: max rel closure {r[wet].max():.3e}", flush=True)
        assert r[wet].max() <= 1e-12


def test_ros2_unsupported_options_are_status_codes(gpu_ready):
    dom = cases.plain_columns(ni=1, nj=2, K=8, seed=0)
    dom.params.adaptive_method = ROS2_ADAPTIVE
    with capi.SedimentColumns.from_domain(dom) as g:
        g.set_bioturbation_zw()
        with pytest.raises(capi.SedcolError) as e:
            g.integrate(0.0, 3600.0)
        assert e.value.code == capi.SEDCOL_EUNSUPPORTED
    p = dom.params.si()
    p["adaptive_method"] = 5
    with capi.SedimentColumns.from_domain(cases.plain_columns(ni=1, nj=2, K=8, seed=0)) as g:
        with pytest.raises(capi.SedcolError) as e:
            g.set_params(p)
        assert e.value.code == capi.SEDCOL_EINVAL
