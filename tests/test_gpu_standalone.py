"""fabmsed1d_hip on the GPU: the standalone column of
examples/standalone/omexdia_p (main.F90), set up from run_sed.nml and
fabm_sed.nml, runs its time loop through the library; the final state equals
the oracle's bitwise and output.dat holds the oracle's profiles at every
output step (same E15.4E3 text)."""
import numpy as np
import pytest

from sedcol.fileio import DENIT_NAME, STATE_NAMES, fortran_e
from sedcol.params import ADAPTIVE_EULER, RUNGE_KUTTA_4, SedParams
from standalone_util import EXE, domain_from_setup, read_dump, read_setup, run, write_namelists
import oracle as O

pytestmark = pytest.mark.gpu


def _run_both(tmp_path, steps, **kw):
    assert EXE.exists(), "fabmsed1d_hip not built (run __graft_entry__.build())"
    write_namelists(tmp_path, **kw)
    r = run(tmp_path, "--setup", "setup.bin")
    assert r.returncode == 0, r.stderr
    s = read_setup(tmp_path / "setup.bin")
    print(f"fabmsed1d_hip: {steps} steps of method {s['ode_method']}, output every {s['output']}", flush=True)
    r = run(tmp_path, "--steps", steps, "--dump", "dump.bin")
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    return s, read_dump(tmp_path / "dump.bin"), (tmp_path / "output.dat").read_text().splitlines()


def _params(s, **kw):
    return SedParams(D0=0.9, dt_min=s["params"]["dt_min"], relative_change_min=s["params"]["relative_change_min"],
                     bcup_dis=s["bcup_dis"], bcup_part=s["bcup_part"], bcdown=3, clamp=1, **kw)


def _check_block(lines, t, s, x):
    """The output block of time t (main.F90:145-156) against state x."""
    i = next(i for i, l in enumerate(lines) if l.split()[1:2] == ["fluxes"] and float(l.split()[0]) == t)
    assert len(lines[i].split()) == 2 + 11   # one record: time, 'fluxes', nvar fluxes
    for k in range(s["knum"]):
        f = lines[i + 1 + k].split()
        assert len(f) == 4 + 11 + 1
        assert f[0] == fortran_e(t, 15, 3).strip()
        assert f[1] == fortran_e(s["zc"][k], 15, 4, 3).strip()
        assert f[2] == fortran_e(s["dz"][k], 15, 4, 3).strip() and f[3] == fortran_e(s["phi"][k], 15, 4, 3).strip()
        assert f[4:15] == [fortran_e(v, 15, 4, 3).strip() for v in x[:, k]], (t, k)


def _denit(s, x):
    p = _params(s).si()
    fT = O.oc_exp(-4500.0 * (1.0 / (s["bdys"][0] + 273.0) - 1.0 / 288.0))
    return np.array([O.kinetics(p, fT, x[:, k])[2] for k in range(s["knum"])]) * 86400.0


def test_reference_namelists_adaptive_euler(gpu_ready, tmp_path):
    """The reference's own run settings: 15 layers, dt 360 s, ADAPTIVE_EULER,
    closed top for solutes (bcup_dissolved_variables = 1), bioturbation
    profile 0; 3000 steps with outputs after 1440 and 2880."""
    s, d, lines = _run_both(tmp_path, 3000)
    assert s["ode_method"] == ADAPTIVE_EULER and s["output"] == 1440
    head = lines[0].split()
    assert head[:4] == ["time(s)", "depth(m)", "layer-height(m)", "porosity()"]
    assert head[4:] == list(STATE_NAMES) + [DENIT_NAME]
    dom = domain_from_setup(s, _params(s))
    orc = O.Oracle(dom, dom.params.si())
    x = dom.conc0.copy()
    for t in (1440, 2880):
        assert orc.run(x, s["dt"], 1440, ADAPTIVE_EULER) == 0
        _check_block(lines, t * s["dt"], s, x)
    assert orc.run(x, s["dt"], 120, ADAPTIVE_EULER) == 0
    assert d["steps"] == 3000
    print("max abs diff", np.max(np.abs(d["conc"] - x)), flush=True)
    assert np.array_equal(d["conc"], x)
    assert np.array_equal(d["fluxes"][5:10], orc.flux[5:10, 0])
    assert np.array_equal(d["fluxes"][:5], s["fluxes"][:5])
    assert np.array_equal(d["denit"], _denit(s, x))
    assert sum(1 for l in lines if " fluxes " in l) == 2


@pytest.mark.parametrize("profile", [1, 3])
def test_rk4_open_top_distributed_flux(gpu_ready, tmp_path, profile):
    """RK4 with concentrations at the top (bcup 2), the distributed-POM flux
    cap (bcup_particulate_variables 4) and bioturbation profile 1 or 3."""
    s, d, lines = _run_both(tmp_path, 130, run={"ode_method": "1", "bcup_dissolved_variables": "2",
                                                 "output": "50"},
                            sed={"bioturbation_profile": str(profile), "distributed_pom_flux": ".true."})
    assert (s["ode_method"], s["bcup_dis"], s["bcup_part"]) == (RUNGE_KUTTA_4, 2, 4)
    dom = domain_from_setup(s, _params(s))
    orc = O.Oracle(dom, dom.params.si(), fcap=s["fcap"], zw=s["zw"] if profile == 3 else None)
    x = dom.conc0.copy()
    for t in (50, 100):
        assert orc.run(x, s["dt"], 50, RUNGE_KUTTA_4) == 0
        _check_block(lines, t * s["dt"], s, x)
    assert orc.run(x, s["dt"], 30, RUNGE_KUTTA_4) == 0
    assert np.array_equal(d["conc"], x)
    assert np.array_equal(d["fluxes"][5:10], orc.flux[5:10, 0])


@pytest.mark.parametrize("method", [0, 3, 4, 5])
def test_other_fixed_step_methods(gpu_ready, tmp_path, method):
    """EULER, RK4-3/8 and the EMP-1/EMP-2 steppers through the same
    namelist key, bitwise like the oracle."""
    s, d, lines = _run_both(tmp_path, 70, run={"ode_method": str(method), "bcup_dissolved_variables": "2",
                                                "output": "25", "dt": "60."})
    assert s["ode_method"] == method and s["dt"] == 60.0
    dom = domain_from_setup(s, _params(s))
    orc = O.Oracle(dom, dom.params.si())
    x = dom.conc0.copy()
    for t in (25, 50):
        assert orc.run(x, s["dt"], 25, method) == 0
        _check_block(lines, t * s["dt"], s, x)
    assert orc.run(x, s["dt"], 20, method) == 0
    assert np.array_equal(d["conc"], x)


@pytest.mark.parametrize("method", [6, 7])
def test_adaptive_emp_and_ros2(gpu_ready, tmp_path, method):
    """ode_method 6 (adaptive EMP) and 7 (ROS2), which the reference's
    ode_solver does not have: each output block is one integrate call over
    output*dt, equal to the oracle's integration over the same spans."""
    s, d, lines = _run_both(tmp_path, 480, run={"ode_method": str(method), "bcup_dissolved_variables": "2",
                                                "output": "240"})
    dom = domain_from_setup(s, _params(s, adaptive_method=method))
    orc = O.Oracle(dom, dom.params.si())
    x = dom.conc0.copy()
    for t in (240, 480):
        assert orc.emp_integrate(x, 240 * s["dt"]) == 0
        _check_block(lines, t * s["dt"], s, x)
    assert np.array_equal(d["conc"], x)
    assert np.all(d["conc"] >= 0.0)
