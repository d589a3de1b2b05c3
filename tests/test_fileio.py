"""Host-side formats and set-up helpers (CPU):
  * Fortran E edit descriptors against hand-worked values;
  * output.dat as the reference's reader (examples/standalone/omexdia_p/
    plotsed1d.py) parses it: header positions of porosity() and the
    denitrification diagnostic, flux lines by their second token, layer
    lines by whitespace;
  * the *_in_soil restart fields (fabm_sediment_component.F90:1362-1463):
    NetCDF round trip, restored / skipped per field as the reference does;
  * update_porosity(from_surface) (fabm_sediment_driver.F90:408-435) by its
    closed forms.
"""
import numpy as np
import pytest

from sedcol import cases, grid
from sedcol.fileio import (DENIT_NAME, STATE_NAMES, OutputDat, apply_restart_fields, fortran_e, load_restart,
                           restart_fields, save_restart, upward_flux_fields)


@pytest.mark.parametrize("x,w,d,e,want", [
    (123.456, 15, 4, 3, "    0.1235E+003"),
    (-0.00123456, 15, 4, 3, "   -0.1235E-002"),
    (0.0, 15, 4, 3, "    0.0000E+000"),
    (1.0e-120, 15, 4, 3, "    0.1000E-119"),
    (0.99996, 15, 4, 3, "    0.1000E+001"),
    (3600.0, 15, 3, None, "      0.360E+04"),
    (31536000.0, 15, 3, None, "      0.315E+08"),
    (2.5e-200, 15, 3, None, "      0.250-199"),
])
def test_fortran_e(x, w, d, e, want):
    assert fortran_e(x, w, d, e) == want


def test_output_dat_parses_like_the_reference_reader(tmp_path):
    K = 5
    g = grid.init_grid_linear(K, 0.004)
    phi = grid.porosity_linear(g["zc"])
    rng = np.random.default_rng(2)
    blocks = []
    path = tmp_path / "output.dat"
    with OutputDat(path) as out:
        for t in (86400.0, 2 * 86400.0):
            fl = rng.standard_normal(11) * 1e-5
            conc = rng.uniform(0, 100, (11, K))
            den = rng.uniform(0, 1, (1, K))
            out.write(t, fl, g["zc"], g["dz"], phi, conc, den)
            blocks.append((t, fl, conc, den))
    lines = open(path).read().splitlines()
    vars_ = lines[0].split()
    assert vars_[:4] == ["time(s)", "depth(m)", "layer-height(m)", "porosity()"]
    assert vars_[4:15] == list(STATE_NAMES)
    idxdenit = vars_.index(DENIT_NAME)
    idxpor = vars_.index("porosity()")
    got = []
    for line in lines[1:]:
        dat = line.split()
        if dat[1] == "fluxes":
            got.append([float(dat[0]), [float(v) for v in dat[2:]], []])
        else:
            got[-1][2].append((float(dat[2]), float(dat[idxpor]), float(dat[idxdenit]),
                               [float(v) for v in dat[4:15]]))
    assert len(got) == 2
    for (t, fl, conc, den), (tg, flg, layers) in zip(blocks, got):
        assert tg == t and np.array_equal(flg, fl)
        assert len(layers) == K
        for k, (dz, por, dn, c) in enumerate(layers):
            assert np.isclose(dz, g["dz"][k], rtol=5e-4) and np.isclose(por, phi[k], rtol=5e-4)
            assert np.isclose(dn, den[0, k], rtol=5e-4)
            assert np.allclose(c, conc[:, k], rtol=5e-4)


def test_restart_fields_round_trip_and_skip_rules(tmp_path):
    rng = np.random.default_rng(1)
    I, J, K = 4, 3, 7
    state = rng.uniform(0, 10, (I, J, K, 11))
    fields = restart_fields(state)
    assert sorted(fields) == sorted(f"{n}_in_soil" for n in STATE_NAMES)
    path = tmp_path / "restart.nc"
    save_restart(path, fields, time_s=86400.0)
    back = load_restart(path)
    assert all(np.array_equal(back[k], fields[k]) for k in fields)
    # a model restarting from it, with one field missing, one of other bounds, one incomplete
    del back["dissolved_oxygen_in_soil"]
    back["dissolved_nitrate_in_soil"] = back["dissolved_nitrate_in_soil"][:, :, :5]
    back["dissolved_ammonium_in_soil"] = back["dissolved_ammonium_in_soil"].copy()
    back["dissolved_ammonium_in_soil"][0, 0, 0] = np.nan
    old = rng.uniform(0, 1, (I, J, K, 11))
    new, report = apply_restart_fields(back, old)
    for i, n in enumerate(STATE_NAMES):
        key = f"{n}_in_soil"
        if n in ("dissolved_oxygen", "dissolved_nitrate", "dissolved_ammonium"):
            assert report[key] != "restored"
            assert np.array_equal(new[..., i], old[..., i])
        else:
            assert report[key] == "restored"
            assert np.array_equal(new[..., i], state[..., i])


def test_update_porosity_from_surface():
    d = cases.plain_columns(ni=3, nj=1, K=12, seed=2)
    off = d.off
    zc = np.concatenate([grid.init_grid_geometric(12, 1e-3, 0.2)["zc"]] * 3)
    surf = np.array([0.81, 0.74, 0.9])
    mask = np.array([0, 1, 0])
    phi, phiI, fc = grid.update_porosity(surf, zc, d.dz, off, porosity_fac=0.9, pom_flux_max=2.0e4, mask=mask)
    for c in (0, 2):
        sl = slice(off[c], off[c + 1])
        z = zc[sl]
        want = surf[c] * (1.0 - 0.9 * (z - z[0]))
        assert phi[sl][0] == surf[c]
        assert np.allclose(phi[sl], want, rtol=1e-15)
        assert phiI[sl][0] == phi[sl][0] and np.allclose(phiI[sl][1:], 0.5 * (phi[sl][1:] + phi[sl][:-1]))
        cap = 2.0e4 / 86400.0 * (1.0 - want) * d.dz[sl]
        assert np.all(fc[sl] <= cap * (1 + 1e-15)) and np.all(np.diff(fc[sl][1:]) <= 0.0)
        assert fc[sl][0] == cap[0] and fc[sl][1] == cap[1]
    sl = slice(off[1], off[2])
    assert np.all(phi[sl] == 1.0) and np.all(phiI[sl] == 1.0)


def test_upward_flux_export_sign_and_layout():
    """*_upward_flux_at_soil_surface = -flux (fabm_sediment_component.F90:
    1798,1812) on the (I, J) grid, column c = i + I j."""
    ni, nj = 3, 2
    flux = np.arange(11 * ni * nj, dtype=np.float64).reshape(11, ni * nj) - 7.0
    f = upward_flux_fields(flux, ni, nj)
    assert set(f) == {n + "_upward_flux_at_soil_surface" for n in STATE_NAMES}
    for v, n in enumerate(STATE_NAMES):
        a = f[n + "_upward_flux_at_soil_surface"]
        assert a.shape == (ni, nj)
        for i in range(ni):
            for j in range(nj):
                assert a[i, j] == -flux[v, i + ni * j]
    with pytest.raises(ValueError):
        upward_flux_fields(flux[:, :5], ni, nj)
