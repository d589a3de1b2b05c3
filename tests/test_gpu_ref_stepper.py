"""The reference's own ode_solver with the GPU right-hand side.

oracle/_ref/ref_gpu_driver is src/utilities/solver_library.F90 of the
reference, compiled unmodified, driving the library's get_rhs through the
ISO_C_BINDING module: its type_sed_gpu extends the reference's
type_rhs_driver (solver_library.F90:37-49) and overrides get_rhs (:48), the
binding a maintainer adds to the reference tree.  Methods 0-3 then run
exactly as the reference's callers run them (fabm_sediment_component.F90:
1715, main.F90:123), one GPU right-hand side per stage.  The result must
equal the oracle's steppers bitwise (which are themselves bitwise equal to
the same reference build driven by the oracle's right-hand side,
test_oracle.py), including ADAPTIVE_EULER's last_min_dt.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from sedcol import cases, capi
from sedcol.params import EULER, RUNGE_KUTTA_4, ADAPTIVE_EULER, RUNGE_KUTTA_4_38, fill_struct
import oracle as O

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "ref_gpu_driver"


def _write_input(path, dom, x, p, method, nsteps, dt):
    K, ncol = int(dom.nlayers[0]), dom.ncol
    cellf = lambda a: np.ascontiguousarray(a.reshape(ncol, K).T)
    with open(path, "wb") as f:
        f.write(np.array([ncol, K, method, nsteps], dtype=np.int32).tobytes())
        f.write(np.array([dt], dtype=np.float64).tobytes())
        # the components only: Fortran stream I/O of a derived type reads no padding
        last = capi.Params._fields_[-1][0]
        f.write(bytes(fill_struct(capi.Params, p))[: getattr(capi.Params, last).offset + 4])
        for a in (dom.dz, dom.dzc, dom.phi, dom.phiI, dom.bfac):
            f.write(cellf(a).tobytes())
        t, c, u = dom.record(0)
        for a in (dom.biot, t, c, u, dom.fortran_state(x)):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())


@pytest.mark.skipif(not EXE.exists(), reason="oracle/_ref/ref_gpu_driver not built (needs the reference tree)")
@pytest.mark.parametrize("method,dt,nsteps", [(EULER, 100.0, 15), (RUNGE_KUTTA_4, 720.0, 25),
                                              (ADAPTIVE_EULER, 7200.0, 3), (RUNGE_KUTTA_4_38, 720.0, 25)])
def test_reference_ode_solver_with_gpu_rhs(gpu_ready, tmp_path, method, dt, nsteps):
    dom = cases.plain_columns(ni=3, nj=1, K=9, seed=8, total=0.08)
    dom.params.dt_min = 10.0
    dom.params.clamp = 0
    p = dom.params.si()
    rng = np.random.default_rng(5)
    x0 = dom.conc0.copy()
    x0[0:2] = rng.uniform(0.0, 3.0e4, size=x0[0:2].shape)
    x0[8] = rng.uniform(0.0, 50.0, size=x0[8].shape)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(fin, dom, x0, p, method, nsteps, dt)
    print(f"ref_gpu_driver method {method}, {nsteps} steps of {dt} s", flush=True)
    r = subprocess.run([str(EXE), str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(fout, dtype=np.float64)
    K, ncol = int(dom.nlayers[0]), dom.ncol
    nx = 11 * K * ncol
    xg = dom.packed_state(raw[:nx])
    lmd = raw[nx]
    xo = x0.copy()
    orc = O.Oracle(dom, p)
    for _ in range(nsteps):
        assert orc.ode_solver(xo, dt, method) == 0
    print("max abs diff", np.max(np.abs(xg - xo)), flush=True)
    assert np.array_equal(xg, xo)
    if method == ADAPTIVE_EULER:
        assert lmd == orc.last_min_dt[0] and lmd < dt
