"""Per-column spin-up (sedcol_spinup) against the oracle run in the same chunks.

The GPU retires a column at the first chunk boundary where its residual
(and, with a flux criterion, its top fluxes) meet the criterion; the oracle
integrates every column chunk by chunk and the expected retirement time and
state of each column are read off its history.  Columns are independent and
the per-column arithmetic is the oracle's, so times and states are compared
bitwise.
"""
import numpy as np
import pytest

from sedcol import capi, spinup
import oracle as O

pytestmark = pytest.mark.gpu

DAY = 86400.0
NCHUNK = 4


@pytest.fixture(scope="module")
def history():
    rng = np.random.Generator(np.random.Philox(5))
    n = 24
    s = spinup.draw_inputs(rng, n)
    dom = spinup.build(s, np.where(np.arange(n) % 3 == 0, 22, 30))
    orc = O.Oracle(dom, dom.params.si())
    x = dom.conc0.copy()
    res, states, fluxes = [], [], []
    for k in range(NCHUNK + 1):
        if k:
            assert orc.emp_integrate(x, DAY, nthreads=8) == 0
        res.append(orc.residual(x, spinup.CFLOOR))
        states.append(x.copy())
        fluxes.append(orc.rhs(x)[1])
    return s, dom, np.array(res), states, np.array(fluxes)


def _expected(first, states, dom):
    off = dom.off
    tc = np.full(dom.ncol, -1.0)
    xe = states[-1].copy()
    for c in range(dom.ncol):
        if first[c] >= 0:
            tc[c] = first[c] * DAY
            xe[:, off[c]:off[c + 1]] = states[first[c]][:, off[c]:off[c + 1]]
    return tc, xe


def _first_true(ok):
    idx = np.argmax(ok, axis=0)
    return np.where(ok.any(axis=0), idx, -1)


def test_spinup_residual_criterion(gpu_ready, history):
    s, dom, res, states, _ = history
    tol = float(np.median(res[NCHUNK]))
    first = _first_true(res < tol)
    assert (first > 0).any() and (first == -1).any()  # some converge in time, some do not
    tc_e, x_e = _expected(first, states, dom)
    with capi.SedimentColumns.from_domain(dom) as g:
        tc, rs, nu = g.spinup(0.0, DAY, NCHUNK * DAY, tol, spinup.CFLOOR)
        xg = g.get_state()
    print("converged at", tc / DAY, flush=True)
    assert np.array_equal(tc, tc_e)
    assert nu == int((first == -1).sum())
    assert np.array_equal(xg, x_e)
    assert np.array_equal(rs, O.Oracle(dom, dom.params.si()).residual(x_e, spinup.CFLOOR))


def test_spinup_flux_criterion(gpu_ready, history):
    s, dom, res, states, fluxes = history
    ref = fluxes[-1]
    floor = spinup.flux_floor(s)
    rtol = 0.1
    tol = 1.0  # residual always met: the flux criterion decides
    sc = np.maximum(np.abs(ref), floor[None, :])
    ok = np.all(np.abs(fluxes - ref[None]) <= rtol * sc[None], axis=1) & (res < tol)
    first = _first_true(ok)
    assert len(set(first.tolist())) > 1
    tc_e, x_e = _expected(first, states, dom)
    with capi.SedimentColumns.from_domain(dom) as g:
        tc, rs, nu = g.spinup(0.0, DAY, NCHUNK * DAY, tol, spinup.CFLOOR, flux_ref=ref, flux_rtol=rtol,
                              flux_floor=floor)
        xg = g.get_state()
    assert np.array_equal(tc, tc_e)
    assert np.array_equal(xg, x_e)


def test_spinup_arguments(gpu_ready, history):
    _, dom, _, _, _ = history
    with capi.SedimentColumns.from_domain(dom) as g:
        for args in ((0.0, 0.0, DAY, 1e-6), (0.0, DAY, -1.0, 1e-6), (0.0, DAY, DAY, 0.0)):
            with pytest.raises(capi.SedcolError) as e:
                g.spinup(*args, spinup.CFLOOR)
            assert e.value.code == capi.SEDCOL_EINVAL
        # t_max = 0: residual only, nothing integrated
        x0 = g.get_state()
        tc, rs, nu = g.spinup(0.0, DAY, 0.0, 1e-30, spinup.CFLOOR)
        assert nu == dom.ncol and np.all(tc == -1.0)
        assert np.array_equal(g.get_state(), x0)


def test_nan_in_a_deep_layer(gpu_ready):
    """A NaN below the top layers reaches the column's residual (the layer
    maximum keeps NaN, as the oracle's scan does) and stops the spin-up with
    SEDCOL_ENAN instead of retiring the column as converged at t = 0."""
    rng = np.random.Generator(np.random.Philox(11))
    n = 6
    s = spinup.draw_inputs(rng, n)
    dom = spinup.build(s, np.full(n, 30))
    x = dom.conc0.copy()
    x[6, dom.off[2] + 5] = np.nan  # NO3, column 2, layer 5
    ro = O.Oracle(dom, dom.params.si()).residual(x, spinup.CFLOOR)
    with capi.SedimentColumns.from_domain(dom) as g:
        g.set_state(x)
        rg = g.residual(spinup.CFLOOR)
        assert np.isnan(ro[2]) and np.isnan(rg[2])
        ok = np.arange(n) != 2
        assert np.array_equal(rg[ok], ro[ok])
        with pytest.raises(capi.SedcolError) as e:
            g.spinup(0.0, DAY, 2 * DAY, 1e30, spinup.CFLOOR)
        assert e.value.code == capi.SEDCOL_ENAN
        assert g.diagnostics()["status"][2] == 1


def test_single_record_entry_points_refuse_a_series(gpu_ready):
    """get_rhs, the residual and the spin-up run under one forcing record;
    with a forcing series active they return SEDCOL_EINVAL."""
    from sedcol import cases

    dom = cases.shelf_sea(n=10, days=3, seed=2)
    assert (dom.mask == 0).any()
    with capi.SedimentColumns.from_domain(dom) as g:
        for call in (lambda: g.get_rhs(dom.conc0), lambda: g.residual(spinup.CFLOOR),
                     lambda: g.spinup(0.0, DAY, DAY, 1e-6, spinup.CFLOOR)):
            with pytest.raises(capi.SedcolError) as e:
                call()
            assert e.value.code == capi.SEDCOL_EINVAL
        g.integrate(0.0, DAY)  # the series itself integrates
    with capi.SedimentColumns.from_domain(dom, record=1, series=False) as g:
        g.get_rhs(dom.conc0)
        g.residual(spinup.CFLOOR)
