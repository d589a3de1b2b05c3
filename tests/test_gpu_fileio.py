"""The library's state through the reference's I/O paths, on the GPU:
  * sedcol_rates (the per-cell DIC and N2 release behind the
    hzg_omexdia_p_denit output) bitwise against the oracle's kinetics;
  * a porosity import mid-run (update_porosity(from_surface),
    fabm_sediment_driver.F90:408-435, then sedcol_set_geometry) continued
    bitwise like the oracle with the same geometry;
  * *_in_soil restart fields written from one handle (NetCDF) and read into
    a fresh one: the continued RK4 run equals the uninterrupted one bitwise;
  * output.dat written from library results.
"""
import numpy as np
import pytest

from sedcol import cases, capi, grid
from sedcol.fileio import (OutputDat, apply_restart_fields, denit_diagnostic, load_restart, restart_fields,
                           save_restart)
from sedcol.params import RUNGE_KUTTA_4
import oracle as O

pytestmark = pytest.mark.gpu


def test_rates_bitwise(gpu_ready):
    dom = cases.plain_columns(ni=3, nj=2, K=17, seed=3)
    rng = np.random.default_rng(3)
    x = dom.conc0.copy()
    x[0:5] = rng.uniform(0, 1e4, x[0:5].shape)
    x[8] = rng.uniform(0, 80, x[8].shape)
    for arith in (0, 1):
        dom.params.arith = arith
        with capi.SedimentColumns.from_domain(dom) as g:
            g.set_state(x)
            dic, n2 = g.rates()
        p = dom.params.si()
        col = np.repeat(np.arange(dom.ncol), dom.nlayers)
        for i in range(dom.ncell):
            fT = O.oc_exp(-4500.0 * (1.0 / (dom.temp[0, col[i]] + 273.0) - 1.0 / 288.0))
            _, sC, sN = O.kinetics(p, fT, x[:, i])
            assert dic[i] == sC and n2[i] == sN, (arith, i)


def test_porosity_import_mid_run_bitwise(gpu_ready):
    dom = cases.plain_columns(ni=2, nj=2, K=20, seed=6)
    dom.params.clamp = 0
    p = dom.params.si()
    zc = np.concatenate([grid.init_grid_geometric(20, 1e-3, 0.2)["zc"]] * dom.ncol)
    phi2, phiI2, _ = grid.update_porosity(np.array([0.86, 0.8, 0.77, 0.83]), zc, dom.dz, dom.off)
    with capi.SedimentColumns.from_domain(dom) as g:
        g.run(360.0, 240, RUNGE_KUTTA_4)
        g.set_geometry(dom.dz, dom.dzc, phi2, phiI2, dom.bfac, dom.biot)
        g.run(360.0, 240, RUNGE_KUTTA_4)
        xg = g.get_state()
    xo = dom.conc0.copy()
    O.Oracle(dom, p).run(xo, 360.0, 240, RUNGE_KUTTA_4)
    dom2 = cases.Domain(**{**dom.__dict__, "phi": phi2, "phiI": phiI2})
    O.Oracle(dom2, p).run(xo, 360.0, 240, RUNGE_KUTTA_4)
    assert np.array_equal(xg, xo)


def test_restart_fields_continue_bitwise(gpu_ready, tmp_path):
    dom = cases.plain_columns(ni=3, nj=2, K=15, seed=9)
    dom.params.clamp = 0
    I, J, K = 3, 2, 15

    def ijkn(x):  # packed (NV, ncell) -> conc(i, j, k, n), column c = i + I j
        return np.ascontiguousarray(x.reshape(11, J, I, K).transpose(2, 1, 3, 0))

    def packed(a):
        return np.ascontiguousarray(a.transpose(3, 1, 0, 2).reshape(11, -1))

    with capi.SedimentColumns.from_domain(dom) as g:
        g.run(360.0, 480, RUNGE_KUTTA_4)
        straight = g.get_state()
    with capi.SedimentColumns.from_domain(dom) as g:
        g.run(360.0, 240, RUNGE_KUTTA_4)
        save_restart(tmp_path / "restart.nc", restart_fields(ijkn(g.get_state())), time_s=86400.0)
    fields = load_restart(tmp_path / "restart.nc")
    state, report = apply_restart_fields(fields, ijkn(dom.conc0))
    assert set(report.values()) == {"restored"}
    with capi.SedimentColumns.from_domain(dom) as g:
        g.set_state(packed(state))
        g.run(360.0, 240, RUNGE_KUTTA_4)
        cont = g.get_state()
    assert np.array_equal(cont, straight)


def test_output_dat_from_the_library(gpu_ready, tmp_path):
    dom = cases.plain_columns(ni=1, nj=1, K=30, seed=0)
    gz = grid.init_grid_geometric(30, 1e-3, 0.2)
    with capi.SedimentColumns.from_domain(dom) as g, OutputDat(tmp_path / "output.dat") as out:
        for day in range(3):
            g.run(360.0, 240, RUNGE_KUTTA_4)
            _, n2 = g.rates()
            out.write((day + 1) * 86400.0, g.fluxes()[:, 0], gz["zc"], dom.dz, dom.phi, g.get_state(),
                      denit_diagnostic(n2)[None, :])
    lines = open(tmp_path / "output.dat").read().splitlines()
    assert len(lines) == 1 + 3 * 31
    assert sum(1 for s in lines if s.split()[1] == "fluxes") == 3
