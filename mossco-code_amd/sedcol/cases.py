"""Synthetic sediment domains of the benchmark configurations.

Every draw comes from one ``numpy.random.Generator(Philox(seed))`` so the
CPU oracle and the GPU library receive bit-identical inputs.  The domains
mirror what the reference's callers assemble before the time loop
(examples/standalone/omexdia_p/main.F90:55-107: grid, porosity,
concentrations, bottom-water ``bdys`` and detritus ``fluxes``), generalised
to per-column porosity, bioturbation and forcing.

Distributions (per column):
  porosity  phi(z) = phi_inf + (phi0 - phi_inf) exp(-z/0.02 m),
            phi0 ~ U[0.75,0.9], phi_inf ~ U[0.55,0.7] (at layer centres);
  bioturbation Db ~ logU[1e-12,1e-10] m2 s-1 (B = Db*86400*1e4 cm2 d-1),
            factor 1 to 5 cm then exp(-(z-0.05)/0.05);
  T ~ U[4,20] degC; bottom water (mmol m-3) O2 U[150,320], NO3 U[2,40],
            NH4 U[0.5,15], PO4 U[0.1,2], ODU 0;
  detritus C flux U[1,50] mmol m-2 d-1, 70/30 labile/semilabile, N and P
            at Redfield C:N:P 106:16:1 (all detritus P into detP);
  initial solutes = bottom water, particulates 0.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import grid
from .params import NVAR, SedParams, PARTICULATE

SECONDS_PER_DAY = 86400.0
SECONDS_PER_YEAR = 365.0 * SECONDS_PER_DAY


@dataclass
class Domain:
    """Packed sediment domain (column c owns cells off[c]..off[c+1]-1)."""
    nlayers: np.ndarray            # (ncol,) int32
    mask: np.ndarray               # (ncol,) int32, 1 = land
    dz: np.ndarray                 # (ncell,)
    dzc: np.ndarray                # (ncell,), 0 at each column's top layer
    phi: np.ndarray                # (ncell,)
    phiI: np.ndarray               # (ncell,)
    bfac: np.ndarray               # (ncell,)
    biot: np.ndarray               # (ncol,) cm2 d-1
    temp: np.ndarray               # (nrec, ncol)
    cbw: np.ndarray                # (nrec, NV, ncol)
    fup: np.ndarray                # (nrec, NV, ncol)
    conc0: np.ndarray              # (NV, ncell)
    params: SedParams = field(default_factory=SedParams)
    t0: float = 0.0
    dt_rec: float = float("inf")
    meta: dict = field(default_factory=dict)

    @property
    def ncol(self) -> int:
        return int(self.nlayers.size)

    @property
    def off(self) -> np.ndarray:
        o = np.zeros(self.ncol + 1, dtype=np.int32)
        np.cumsum(self.nlayers, out=o[1:])
        return o

    @property
    def ncell(self) -> int:
        return int(self.nlayers.sum())

    @property
    def nrec(self) -> int:
        return int(self.temp.shape[0])

    @property
    def kmax(self) -> int:
        return int(self.nlayers.max())

    @property
    def uniform(self) -> bool:
        return bool(np.all(self.nlayers == self.nlayers[0]))

    def record(self, r: int):
        """(temp, cbw, fup) of forcing record r, contiguous.  fup may hold a
        single record that applies to the whole series."""
        rf = r if self.fup.shape[0] > 1 else 0
        return (np.ascontiguousarray(self.temp[r]), np.ascontiguousarray(self.cbw[r]),
                np.ascontiguousarray(self.fup[rf]))

    def subset(self, cols) -> "Domain":
        """Domain made of the given columns (in that order)."""
        cols = np.asarray(cols, dtype=np.int64)
        off = self.off
        cell = np.concatenate([np.arange(off[c], off[c + 1]) for c in cols]) if cols.size else np.zeros(0, np.int64)
        return Domain(
            nlayers=self.nlayers[cols].copy(), mask=self.mask[cols].copy(),
            dz=self.dz[cell].copy(), dzc=self.dzc[cell].copy(), phi=self.phi[cell].copy(),
            phiI=self.phiI[cell].copy(), bfac=self.bfac[cell].copy(), biot=self.biot[cols].copy(),
            temp=self.temp[:, cols].copy(), cbw=self.cbw[:, :, cols].copy(), fup=self.fup[:, :, cols].copy(),
            conc0=self.conc0[:, cell].copy(), params=self.params, t0=self.t0, dt_rec=self.dt_rec,
            meta=dict(self.meta, subset=True))

    def fortran_state(self, x: np.ndarray) -> np.ndarray:
        """Packed (NV, ncell) -> Fortran (NV, K, ncol) for a uniform grid."""
        K = int(self.nlayers[0])
        return np.ascontiguousarray(x.reshape(NVAR, self.ncol, K).transpose(0, 2, 1))

    def packed_state(self, xf: np.ndarray) -> np.ndarray:
        K = int(self.nlayers[0])
        return np.ascontiguousarray(xf.reshape(NVAR, K, self.ncol).transpose(0, 2, 1).reshape(NVAR, -1))


# ---------------------------------------------------------------------------
# column scalar draws
# ---------------------------------------------------------------------------

SCALAR_RANGES = {
    "phi0": (0.75, 0.9), "phi_inf": (0.55, 0.7), "log10_Db": (-12.0, -10.0), "T": (4.0, 20.0),
    "O2": (150.0, 320.0), "NO3": (2.0, 40.0), "NH4": (0.5, 15.0), "PO4": (0.1, 2.0), "Cflux": (1.0, 50.0),
}
SCALAR_NAMES = tuple(SCALAR_RANGES)


def draw_scalars(rng: np.random.Generator, ncol: int) -> dict:
    """Independent uniform draws in SCALAR_RANGES (log10_Db uniform = Db
    log-uniform)."""
    out = {}
    for name in SCALAR_NAMES:
        lo, hi = SCALAR_RANGES[name]
        out[name] = lo + (hi - lo) * rng.random(ncol)
    return out


def scalars_from_unit(u: dict) -> dict:
    """Map unit-interval fields to SCALAR_RANGES."""
    return {n: SCALAR_RANGES[n][0] + (SCALAR_RANGES[n][1] - SCALAR_RANGES[n][0]) * u[n] for n in SCALAR_NAMES}


def bottom_water(s: dict) -> np.ndarray:
    """(NV, ncol) bottom-water concentrations; particulates and ODU 0."""
    ncol = s["T"].size
    cbw = np.zeros((NVAR, ncol))
    cbw[5] = s["O2"]
    cbw[6] = s["NO3"]
    cbw[7] = s["NH4"]
    cbw[9] = s["PO4"]
    return cbw


def deposition(s: dict, frac_labile=0.7) -> np.ndarray:
    """(NV, ncol) top fluxes (mmol m-2 s-1, positive downward)."""
    F = s["Cflux"] / SECONDS_PER_DAY
    ncol = F.size
    fup = np.zeros((NVAR, ncol))
    fup[0] = frac_labile * F
    fup[1] = (1.0 - frac_labile) * F
    fup[2] = frac_labile * F * 16.0 / 106.0
    fup[3] = (1.0 - frac_labile) * F * 16.0 / 106.0
    fup[4] = F / 106.0
    return fup


#: domains of at least this many columns build their geometry with array
#: operations (the same formulas; numpy's power may round differently in the
#: last bit from the per-column path)
LARGE = 20000


def _geometric_cells(nlayers, total_depth, dz1, off, colid):
    """Cell geometry of geometric columns, vectorised over all cells."""
    r = grid.geometric_ratios(nlayers, dz1, total_depth)
    first = off[:-1]
    kidx = np.arange(colid.size) - np.repeat(first, nlayers)
    dz = dz1 * r[colid] ** kidx
    # interface depths: running sums of dz within each column, top first
    # (columns of one layer count at a time, as (n, K) rows)
    zi = np.empty_like(dz)
    for K in np.unique(nlayers):
        cols = np.flatnonzero(nlayers == K)
        cells = (first[cols][:, None] + np.arange(K)[None, :])
        d = dz[cells]
        z = np.zeros_like(d)
        if K > 1:
            z[:, 1:] = np.cumsum(d[:, :-1], axis=1)
        zi[cells] = z
    zc = zi + 0.5 * dz
    dzc = np.empty_like(dz)
    dzc[1:] = zc[1:] - zc[:-1]
    dzc[first] = 0.0
    return dz, dzc, zc, zi


def build_columns(nlayers, total_depth, s: dict, dz1=1.0e-3, mask=None, params=None, meta=None) -> Domain:
    """Geometry and initial state of columns with geometric layers."""
    nlayers = np.asarray(nlayers, dtype=np.int32)
    ncol = nlayers.size
    total_depth = np.broadcast_to(np.asarray(total_depth, dtype=np.float64), (ncol,))
    off = np.zeros(ncol + 1, dtype=np.int64)
    np.cumsum(nlayers, out=off[1:])
    ncell = int(off[-1])
    colid = np.repeat(np.arange(ncol), nlayers)
    if ncol >= LARGE:
        dz, dzc, zc, zi = _geometric_cells(nlayers, total_depth, dz1, off, colid)
    else:
        dz = np.empty(ncell)
        dzc = np.empty(ncell)
        zc = np.empty(ncell)
        zi = np.empty(ncell)
        cache = {}
        for c in range(ncol):
            key = (int(nlayers[c]), float(total_depth[c]))
            g = cache.get(key)
            if g is None:
                g = grid.init_grid_geometric(key[0], dz1, key[1])
                cache[key] = g
            sl = slice(off[c], off[c + 1])
            dz[sl] = g["dz"]
            dzc[sl] = g["dzc"]
            zc[sl] = g["zc"]
            zi[sl] = g["zi"][:-1]
    phi = grid.porosity_exponential(zc, s["phi0"][colid], s["phi_inf"][colid])
    phiI = np.empty(ncell)
    first = off[:-1]
    phiI[:] = 0.5 * (np.roll(phi, 1) + phi)
    phiI[first] = phi[first]
    bfac = grid.bioturbation_const_exp(zi)
    Db = 10.0 ** s["log10_Db"]
    biot = Db * SECONDS_PER_DAY * 1.0e4
    cbw = bottom_water(s)
    fup = deposition(s)
    conc0 = np.zeros((NVAR, ncell))
    for n in range(NVAR):
        if not PARTICULATE[n]:
            conc0[n] = cbw[n][colid]
    m = np.zeros(ncol, dtype=np.int32) if mask is None else np.asarray(mask, dtype=np.int32)
    for n in range(NVAR):
        conc0[n][m[colid] != 0] = 1.0e20  # missing value of land cells (fabm_sediment_driver.F90:86,459-468)
    return Domain(nlayers=nlayers, mask=m, dz=dz, dzc=dzc, phi=phi, phiI=phiI, bfac=bfac, biot=biot,
                  temp=s["T"][None, :].copy(), cbw=cbw[None].copy(), fup=fup[None].copy(), conc0=conc0,
                  params=params or SedParams(), meta=dict(meta or {}, scalars=s))


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------

def plain_columns(ni=4, nj=4, K=30, seed=0, total=0.2) -> Domain:
    """Plain RK4 case: ni x nj columns, K geometric layers (1 mm top, 0.2 m),
    RK4 with dt = 360 s for 30 days."""
    rng = np.random.Generator(np.random.Philox(seed))
    ncol = ni * nj
    s = draw_scalars(rng, ncol)
    p = SedParams(bcup_dis=2, bcup_part=1, bcdown=3, clamp=1)
    d = build_columns(np.full(ncol, K), total, s, params=p,
                      meta={"name": "plain-rk4", "ni": ni, "nj": nj, "seed": seed,
                            "method": "rk4", "dt": 360.0, "days": 30})
    return d


def _smooth_unit_fields(rng, shape, names, sigma):
    """Gaussian-smoothed noise mapped to U(0,1) marginals."""
    from scipy.ndimage import gaussian_filter
    from scipy.special import ndtr
    out = {}
    for n in names:
        f = gaussian_filter(rng.standard_normal(shape), sigma=sigma, mode="wrap")
        f = (f - f.mean()) / f.std()
        out[n] = ndtr(f).ravel()
    return out


def shelf_sea(n=300, K=30, land_fraction=0.4, sigma=10.0, days=365, seed=0, total=0.2,
              seasonal_amp=0.3, temp_amp=5.0, noise=0.05) -> Domain:
    """Shelf-sea case: n x n columns, random-blob land mask, spatially
    smoothed inputs, daily bottom-water forcing for `days` days:
      c_d = c_mean (1 + a sin(2 pi d/365 + phase) + noise*eps), a = 0.3;
      T_d = T_mean + 5 sin(2 pi d/365 + phase) + 0.5 eps, phase = -pi/2
      (coldest on day 0), O2 in anti-phase with T; eps ~ N(0,1) per
      column and day.  Detritus deposition is constant per column."""
    rng = np.random.Generator(np.random.Philox(seed))
    ncol = n * n
    u = _smooth_unit_fields(rng, (n, n), SCALAR_NAMES + ("land",), sigma)
    land = (u.pop("land") < land_fraction).astype(np.int32)
    s = scalars_from_unit(u)
    p = SedParams(bcup_dis=2, bcup_part=1, bcdown=3, clamp=0, emp_rtol=1e-6, emp_atol=1e-6,
                  emp_dt_min=1.0, emp_dt_max=3600.0)
    d = build_columns(np.full(ncol, K), total, s, mask=land, params=p,
                      meta={"name": "shelf-emp", "n": n, "seed": seed, "days": days, "land_fraction": land_fraction,
                            "sigma": sigma, "seasonal_amp": seasonal_amp, "temp_amp": temp_amp, "noise": noise})
    day = np.arange(days, dtype=np.float64)
    phase = -0.5 * np.pi
    season = np.sin(2.0 * np.pi * day / 365.0 + phase)[:, None]
    eps_T = rng.standard_normal((days, ncol))
    temp = s["T"][None, :] + temp_amp * season + 0.5 * eps_T
    cbw = np.repeat(d.cbw, days, axis=0)
    for v, sign in ((5, -1.0), (6, 1.0), (7, 1.0), (9, 1.0)):
        eps = rng.standard_normal((days, ncol))
        cbw[:, v, :] = np.maximum(s[("O2", "NO3", "NH4", None, "PO4")[v - 5]][None, :]
                                  * (1.0 + sign * seasonal_amp * season + noise * eps), 0.0)
    d.temp = temp
    d.cbw = cbw
    d.t0 = 0.0
    d.dt_rec = SECONDS_PER_DAY
    return d


def ragged_columns(ncol=100_000, kmin=10, kmax=60, depth=(0.1, 0.5), seed=0) -> Domain:
    """Ragged batch: K ~ U{kmin..kmax}, total depth ~ U[depth], geometric
    layers with a 1 mm top layer; other inputs as in plain_columns."""
    rng = np.random.Generator(np.random.Philox(seed))
    K = rng.integers(kmin, kmax + 1, size=ncol).astype(np.int32)
    total = depth[0] + (depth[1] - depth[0]) * rng.random(ncol)
    s = draw_scalars(rng, ncol)
    p = SedParams(bcup_dis=2, bcup_part=1, bcdown=3, clamp=0, emp_rtol=1e-6, emp_atol=1e-6,
                  emp_dt_min=1.0, emp_dt_max=3600.0)
    return build_columns(K, total, s, params=p, meta={"name": "ragged-emp", "seed": seed})


def record_spans(t_start: float, t_end: float, nrec: int, t0: float, dt_rec: float):
    """(record, span) pairs covering [t_start, t_end) with piecewise-constant
    records; the same arithmetic as the adaptive kernel."""
    import math

    def rec_of(t):
        if nrec <= 1:
            return 0
        q = math.floor((t - t0) / dt_rec)
        if not q > 0:
            return 0
        return nrec - 1 if q >= nrec - 1 else int(q)

    out = []
    for r in range(rec_of(t_start), rec_of(t_end) + 1):
        lo = -math.inf if r == 0 else t0 + float(r) * dt_rec
        hi = math.inf if r == nrec - 1 else t0 + float(r + 1) * dt_rec
        a = t_start if t_start > lo else lo
        b = t_end if t_end < hi else hi
        out.append((r, b - a))
    return out
