"""Steady states by per-column spin-up, and the surrogate's data formats.

The reference spins the sediment up by integrating one column with constant
bottom-water conditions for `presimulation_years` (dt = 3600 s) and copying
it to every wet column (fabm_sediment_component.F90:557-632).  Here every
column is spun up on its own inputs (sedcol_spinup: adaptive EMP, chunked,
columns retired once their steady-state residual, and optionally their top
fluxes, meet the criterion).  The same machinery produces the surrogate's
training pairs and the cold-start reference it is judged against.

Inputs (12 scalars per column, in this order):
  T (degC), bottom-water O2, NO3, NH4, PO4, ODU (mmol m-3),
  log10 of the detritus carbon deposition (mmol C m-2 d-1), labile fraction
  of the deposition, surface and deep porosity phi0 / phi_inf,
  log10 bioturbation diffusivity (m2 s-1), total sediment depth (m).
Targets: log10(c + eps_n) of the 11 state variables on the 30 geometric
layers (dz1 = 1 mm) of the column's own depth, 330 values, variable-major.
"""
from __future__ import annotations

import numpy as np

from . import cases, grid
from .params import NVAR, PARTICULATE, SedParams

DAY = 86400.0
YEAR = 365.0 * DAY
K_SURR = 30

INPUTS = ("T", "O2", "NO3", "NH4", "PO4", "ODU", "log10_Cflux", "frac_labile", "phi0", "phi_inf", "log10_Db",
          "depth")
#: training ranges (uniform); the first nine follow the plain-column draws
RANGES = {
    "T": (4.0, 20.0), "O2": (150.0, 320.0), "NO3": (2.0, 40.0), "NH4": (0.5, 15.0), "PO4": (0.1, 2.0),
    "ODU": (0.0, 10.0), "log10_Cflux": (0.0, float(np.log10(50.0))), "frac_labile": (0.6, 0.8),
    "phi0": (0.75, 0.9), "phi_inf": (0.55, 0.7), "log10_Db": (-12.0, -10.0), "depth": (0.15, 0.3),
}
#: physical limits for draws outside the training ranges
LIMITS = {
    "T": (-1.0, 30.0), "O2": (10.0, 400.0), "NO3": (0.0, 60.0), "NH4": (0.0, 25.0), "PO4": (0.0, 3.0),
    "ODU": (0.0, 20.0), "log10_Cflux": (-0.5, 2.0), "frac_labile": (0.4, 0.95), "phi0": (0.6, 0.95),
    "phi_inf": (0.4, 0.75), "log10_Db": (-13.0, -9.5), "depth": (0.1, 0.45),
}
#: residual floor per variable (mmol m-3) for the steady-state residual
CFLOOR = np.full(NVAR, 1.0e-3)
#: offsets of the log10 targets (mmol m-3)
EPS = np.full(NVAR, 1.0e-3)


def draw_inputs(rng: np.random.Generator, n: int, tails: float = 0.0, tail_width: float = 0.25) -> dict:
    """n input sets uniform in RANGES; a fraction `tails` of them (the last
    ones after a shuffle) has each input, with probability 1/2, drawn
    outside its training range by up to tail_width of the range's width
    (clipped to LIMITS).  Returns the inputs and a boolean 'tail' array."""
    s = {}
    for k in INPUTS:
        lo, hi = RANGES[k]
        s[k] = lo + (hi - lo) * rng.random(n)
    ntail = int(round(tails * n))
    tail = np.zeros(n, dtype=bool)
    if ntail:
        idx = rng.permutation(n)[:ntail]
        tail[idx] = True
        for k in INPUTS:
            lo, hi = RANGES[k]
            w = (hi - lo) * tail_width
            pick = idx[rng.random(ntail) < 0.5]
            side = rng.random(pick.size) < 0.5
            off = w * rng.random(pick.size)
            v = np.where(side, lo - off, hi + off)
            s[k][pick] = np.clip(v, *LIMITS[k])
    s["tail"] = tail
    return s


def as_matrix(s: dict) -> np.ndarray:
    """(n, 12) float64 inputs in INPUTS order."""
    return np.stack([np.asarray(s[k], dtype=np.float64) for k in INPUTS], axis=1)


def build(s: dict, nlayers, params: SedParams | None = None) -> cases.Domain:
    """Columns (geometric layers, 1 mm top, the drawn depth) with constant
    forcing from the inputs and the cold initial state (solutes at bottom
    water, no detritus)."""
    n = np.asarray(s["T"]).size
    nlayers = np.broadcast_to(np.asarray(nlayers, dtype=np.int32), (n,)).copy()
    sc = {
        "T": s["T"], "O2": s["O2"], "NO3": s["NO3"], "NH4": s["NH4"], "PO4": s["PO4"],
        "Cflux": 10.0 ** np.asarray(s["log10_Cflux"]), "phi0": s["phi0"], "phi_inf": s["phi_inf"],
        "log10_Db": s["log10_Db"],
    }
    p = params or SedParams()
    dom = cases.build_columns(nlayers, s["depth"], sc, params=p, meta={"name": "spinup"})
    # ODU bottom water and the labile fraction of the deposition
    dom.cbw[0, 8] = s["ODU"]
    F = sc["Cflux"] / DAY
    fl = np.asarray(s["frac_labile"], dtype=np.float64)
    dom.fup[0, 0] = fl * F
    dom.fup[0, 1] = (1.0 - fl) * F
    dom.fup[0, 2] = fl * F * 16.0 / 106.0
    dom.fup[0, 3] = (1.0 - fl) * F * 16.0 / 106.0
    dom.fup[0, 4] = F / 106.0
    colid = np.repeat(np.arange(n), nlayers)
    dom.conc0[8] = np.asarray(s["ODU"])[colid]
    return dom


def flux_floor(s: dict) -> np.ndarray:
    """Per-column absolute floor of the flux criterion: 1 % of the carbon
    deposition (mmol m-2 s-1)."""
    return 0.01 * 10.0 ** np.asarray(s["log10_Cflux"]) / DAY


def to_targets(x: np.ndarray, nlayers: int = K_SURR) -> np.ndarray:
    """Packed state (NV, n*K) of uniform-K columns -> (n, NV*K) log10 targets."""
    n = x.shape[1] // nlayers
    y = np.log10(np.maximum(x, 0.0) + EPS[:, None])
    return np.ascontiguousarray(y.reshape(NVAR, n, nlayers).transpose(1, 0, 2).reshape(n, NVAR * nlayers))


def layer_centres(nlayers: int, depth: float, dz1: float = 1.0e-3) -> np.ndarray:
    return grid.init_grid_geometric(int(nlayers), dz1, float(depth))["zc"]


def _centres_many(nlayers: int, depth: np.ndarray, dz1: float = 1.0e-3) -> np.ndarray:
    """(n, K) layer centres of geometric columns of the given depths."""
    r = grid.geometric_ratios(np.full(depth.size, nlayers), dz1, depth)
    dz = dz1 * r[:, None] ** np.arange(nlayers)[None, :]
    zi = np.zeros_like(dz)
    zi[:, 1:] = np.cumsum(dz[:, :-1], axis=1)
    return zi + 0.5 * dz


def _interp_rows(zk: np.ndarray, z30: np.ndarray, f: np.ndarray) -> np.ndarray:
    """Row-wise linear interpolation, constant beyond the end points:
    zk (n, K) targets, z30 (n, M) increasing nodes, f (n, V, M) -> (n, V, K)."""
    n, M = z30.shape
    j = np.empty(zk.shape, dtype=np.int64)
    for i0 in range(0, n, 65536):  # bounded temporaries
        i1 = min(n, i0 + 65536)
        j[i0:i1] = (zk[i0:i1, :, None] >= z30[i0:i1, None, :]).sum(axis=2) - 1
    j = np.clip(j, 0, M - 2)
    rows = np.arange(n)[:, None]
    z0, z1 = z30[rows, j], z30[rows, j + 1]
    w = np.clip((zk - z0) / (z1 - z0), 0.0, 1.0)
    f0 = np.take_along_axis(f, j[:, None, :], axis=2)
    f1 = np.take_along_axis(f, j[:, None, :] + 1, axis=2)
    return f0 + w[:, None, :] * (f1 - f0)


def from_targets(y: np.ndarray, depth, nlayers) -> np.ndarray:
    """(n, NV*30) log10 profiles on each column's 30-layer grid -> packed
    state (NV, sum K) on K-layer grids of the same depth: linear
    interpolation of the log10 profile in depth (constant beyond the first
    and last centres), then 10**y - eps clipped at 0.  Large batches are
    interpolated with array operations, one layer count at a time."""
    y = np.asarray(y, dtype=np.float64)
    n = y.shape[0]
    depth = np.broadcast_to(np.asarray(depth, dtype=np.float64), (n,))
    nlayers = np.broadcast_to(np.asarray(nlayers, dtype=np.int64), (n,))
    prof = y.reshape(n, NVAR, K_SURR)
    if n >= 2048:
        off = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(nlayers, out=off[1:])
        out = np.empty((NVAR, int(off[-1])))
        z30 = _centres_many(K_SURR, np.ascontiguousarray(depth))
        for K in np.unique(nlayers):
            cols = np.flatnonzero(nlayers == K)
            zk = _centres_many(int(K), np.ascontiguousarray(depth[cols]))
            v = _interp_rows(zk, z30[cols], prof[cols])  # (nc, NV, K)
            cells = off[cols][:, None] + np.arange(K)[None, :]
            out[:, cells.ravel()] = v.transpose(1, 0, 2).reshape(NVAR, -1)
        return np.maximum(10.0 ** out - EPS[:, None], 0.0)
    out = np.empty((NVAR, int(nlayers.sum())))
    o = 0
    cache = {}
    for c in range(n):
        key = (float(depth[c]), int(nlayers[c]))
        zz = cache.get(key)
        if zz is None:
            zz = (layer_centres(K_SURR, key[0]), layer_centres(key[1], key[0]))
            cache[key] = zz
        z30, zk = zz
        K = key[1]
        for v in range(NVAR):
            out[v, o:o + K] = np.interp(zk, z30, prof[c, v])
        o += K
    return np.maximum(10.0 ** out - EPS[:, None], 0.0)


def spin(dom: cases.Domain, years: float, tol: float, chunk_days: float = 10.0, conc=None, flux_ref=None,
         flux_rtol: float = 0.01, floor=None, columns=None, segment_years: float = 0.25, log=None, init=None):
    """Spin the domain's columns up on the GPU from `conc` (default: the
    cold state), or from the state that init(handle) sets on the device (a
    warm start, SedimentColumns.set_state_profiles), in segments of
    segment_years (a progress line each through `log`).  Returns
    dict(state, t_conv, resid, n_unconverged, flux, wall_s, init_s), init_s
    the time of init."""
    import time

    from . import capi

    g = columns or capi.SedimentColumns.from_domain(dom)
    try:
        if conc is not None:
            g.set_state(conc)
        t_init = time.perf_counter()
        if init is not None:
            init(g)
        t_init = time.perf_counter() - t_init
        d0 = g.diagnostics()
        t0 = time.perf_counter()
        t, t_end = 0.0, years * YEAR
        tconv = np.full(dom.ncol, -1.0)
        while True:
            seg = min(segment_years * YEAR, t_end - t)
            tc, rs, nu = g.spinup(t, chunk_days * DAY, seg, tol, CFLOOR, flux_ref=flux_ref, flux_rtol=flux_rtol,
                                  flux_floor=floor)
            new = (tc >= 0.0) & (tconv < 0.0)
            tconv[new] = t + tc[new]
            t += seg
            if log:
                log(f"spin-up t={t / YEAR:.3f} y unconverged {nu} resid q50 {np.median(rs):.2e} "
                    f"max {rs.max():.2e} wall {time.perf_counter() - t0:.1f} s")
            if nu == 0 or not t < t_end:
                break
        wall = time.perf_counter() - t0
        d1 = g.diagnostics()
        x = g.get_state()
        _, fl = g.get_rhs(x)
    finally:
        if columns is None:
            g.close()
    att = (d1["accepted"] - d0["accepted"]) + (d1["rejected"] - d0["rejected"])
    return {"state": x, "t_conv": tconv, "resid": rs, "n_unconverged": nu, "flux": fl, "wall_s": wall,
            "attempts": att, "init_s": t_init}


def generate(n: int, seed: int, years: float, tol: float = 1.0e-8, nlayers=K_SURR, tails: float = 0.0,
             chunk_days: float = 10.0, method: int = 0, dt_max: float | None = None):
    """Cold spin-up of n drawn columns (nlayers: int or per-column array)
    with adaptive stepper `method` (0 or 6 adaptive EMP, 7 linearly
    implicit) and steps up to dt_max (None: the parameter default).
    Returns the inputs, the final packed state, per-column residual and
    convergence time, top fluxes at the final state and the wall time."""
    rng = np.random.Generator(np.random.Philox(seed))
    s = draw_inputs(rng, n, tails=tails)
    if np.isscalar(nlayers) and nlayers == 0:
        nlayers = rng.integers(20, 41, size=n)
    dom = build(s, nlayers)
    dom.params.adaptive_method = method
    if dt_max is not None:
        dom.params.emp_dt_max = dt_max
    out = spin(dom, years, tol, chunk_days=chunk_days, log=lambda m: print(m, flush=True))
    out["inputs"] = s
    out["nlayers"] = np.asarray(dom.nlayers)
    return dom, out


def _main(argv=None):
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(description="cold spin-up of drawn columns (surrogate training / reference data)")
    ap.add_argument("--n", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--years", type=float, default=2.0)
    ap.add_argument("--tol", type=float, default=1.0e-8)
    ap.add_argument("--ragged", action="store_true", help="layers U{20..40} instead of 30")
    ap.add_argument("--tails", type=float, default=0.0)
    ap.add_argument("--method", type=int, default=0, help="adaptive stepper: 0 EMP, 7 linearly implicit ROS2")
    ap.add_argument("--dt-max", type=float, default=None, help="largest step (s); default: the parameter default")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    t = time.perf_counter()
    dom, o = generate(a.n, a.seed, a.years, a.tol, nlayers=0 if a.ragged else K_SURR, tails=a.tails,
                      method=a.method, dt_max=a.dt_max)
    s = o["inputs"]
    save = {"X": as_matrix(s), "tail": s["tail"], "nlayers": o["nlayers"], "resid": o["resid"],
            "t_conv": o["t_conv"], "flux": o["flux"], "wall_s": o["wall_s"]}
    if a.ragged:
        save["state"] = o["state"].astype(np.float64)
    else:
        save["Y"] = to_targets(o["state"]).astype(np.float32)
    np.savez_compressed(a.out, **save)
    r = o["resid"]
    print(json.dumps({"n": a.n, "years": a.years, "wall_s": o["wall_s"], "total_s": time.perf_counter() - t,
                      "resid_q50": float(np.median(r)), "resid_q90": float(np.quantile(r, 0.9)),
                      "frac_below_1e-6": float(np.mean(r < 1e-6)), "frac_below_1e-8": float(np.mean(r < 1e-8)),
                      "unconverged": o["n_unconverged"]}), flush=True)


if __name__ == "__main__":
    _main()
