"""Write tests/golden/ref_steppers.npz from the reference time steppers.

The vectors come from src/utilities/solver_library.F90 of the reference,
compiled unmodified into oracle/_ref/libref_solver.so (oracle/Makefile, target
``ref``) and driven by the oracle's right-hand side.  They pin the oracle's
EULER, RUNGE_KUTTA_4, ADAPTIVE_EULER and RUNGE_KUTTA_4_38 stepping where the
reference build is absent.  Run from the repository root after
``make -C oracle ref``:

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO / "mossco-code_amd"))
sys.path.insert(0, str(REPO / "oracle"))

from sedcol import cases  # noqa: E402
import oracle as O  # noqa: E402

CASES = [
    # name, method, dt, nsteps
    ("euler", 0, 60.0, 20),
    ("rk4", 1, 360.0, 40),
    ("adaptive_euler", 2, 3600.0, 4),
    ("rk4_38", 3, 360.0, 40),
]


def domain():
    d = cases.plain_columns(ni=2, nj=1, K=6, seed=21, total=0.05)
    d.params.dt_min = 1.0
    d.params.clamp = 0
    return d


def initial_state(d):
    rng = np.random.default_rng(2024)
    x = d.conc0.copy()
    x[0:2] = rng.uniform(1.0e3, 4.0e4, size=x[0:2].shape)
    x[2:4] = x[0:2] * 0.15
    x[4] = x[0] / 106.0
    x[8] = rng.uniform(0.0, 80.0, size=x[8].shape)
    x[10] = rng.uniform(0.0, 5.0, size=x[10].shape)
    return x


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref_solver.so missing: run make -C oracle ref")
    d = domain()
    p = d.params.si()
    x0 = initial_state(d)
    out = {"x0": x0}
    for name, method, dt, nsteps in CASES:
        orc = O.Oracle(d, p)
        x = x0.copy()
        orc.ref_ode_solver(x, dt, method, nsteps=nsteps)
        out[name] = x
        out[name + "_last_min_dt"] = np.array([orc.last_min_dt[0]])
    np.savez(Path(__file__).with_name("ref_steppers.npz"), **out)
    print("wrote", Path(__file__).with_name("ref_steppers.npz"))


if __name__ == "__main__":
    main()
