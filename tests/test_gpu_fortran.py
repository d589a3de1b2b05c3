"""The Fortran host (mossco-code_amd/fortran/fabmsed_hip, built on the
sedcol_hip ISO_C_BINDING module) drives the GPU through the C ABI with the
reference's call sequence; its result must equal the oracle's bitwise."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from sedcol import cases
from sedcol.params import SedParams, RUNGE_KUTTA_4
import oracle as O

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parents[1] / "mossco-code_amd" / "fortran" / "fabmsed_hip"


def read_output(path):
    raw = open(path, "rb").read()
    inum, jnum, knum, nsteps = np.frombuffer(raw[:16], dtype=np.int32)
    a = np.frombuffer(raw[16:], dtype=np.float64)
    nc = inum * jnum
    sizes = [("dz", nc * knum), ("dzc", nc * knum), ("phi", nc * knum), ("phiI", nc * knum), ("bfac", nc * knum),
             ("conc0", nc * knum * 11), ("bdys", nc * 12), ("flux0", nc * 11), ("conc", nc * knum * 11),
             ("flux", nc * 11)]
    out, o = {}, 0
    for name, n in sizes:
        out[name] = a[o:o + n]
        o += n
    assert o == a.size
    return int(inum), int(jnum), int(knum), int(nsteps), out


def test_fortran_host_matches_oracle(gpu_ready, tmp_path):
    assert EXE.exists(), "fabmsed_hip not built (run __graft_entry__.build())"
    out = tmp_path / "run.bin"
    r = subprocess.run([str(EXE), str(out), "240", "360.0"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    inum, jnum, knum, nsteps, f = read_output(out)
    nc = inum * jnum
    cellp = lambda a: np.ascontiguousarray(a.reshape(knum, nc).T).ravel()  # (K, ncol) Fortran -> packed
    statep = lambda a: np.ascontiguousarray(a.reshape(11, knum, nc).transpose(0, 2, 1)).reshape(11, -1)
    bdys = f["bdys"].reshape(12, nc)
    dom = cases.Domain(
        nlayers=np.full(nc, knum, dtype=np.int32), mask=np.zeros(nc, dtype=np.int32),
        dz=cellp(f["dz"]), dzc=cellp(f["dzc"]), phi=cellp(f["phi"]), phiI=cellp(f["phiI"]),
        bfac=cellp(f["bfac"]), biot=np.full(nc, 4.5),
        temp=bdys[0][None].copy(), cbw=bdys[1:][None].copy(), fup=f["flux0"].reshape(1, 11, nc).copy(),
        conc0=statep(f["conc0"]), params=SedParams(bcup_dis=2, bcup_part=1, bcdown=3, clamp=1))
    orc = O.Oracle(dom, dom.params.si())
    x = dom.conc0.copy()
    assert orc.run(x, 360.0, nsteps, RUNGE_KUTTA_4, nthreads=4) == 0
    xg = statep(f["conc"])
    print("max abs diff", np.max(np.abs(xg - x)), flush=True)
    assert np.array_equal(xg, x)
    # solute top fluxes of the last right-hand side, as fluxes(:,:,n) = intFlux(:,:,1)
    fl = f["flux"].reshape(11, nc)
    assert np.array_equal(fl[5:10], orc.flux[5:10])
    assert np.array_equal(fl[0:5], dom.fup[0, 0:5])
