"""Sediment grid, porosity and bioturbation profiles (host set-up).

Mirrors the set-up half of the reference driver, which runs once per
domain on the host and whose results the library receives through
``sedcol_set_geometry``:

* ``init_grid``           src/drivers/fabm_sediment_driver.F90:127-177
* porosity / bioturbation in ``initialize`` :278-303
* ``update_porosity``     :393-442 (from the surface value :408-428, interface
  porosity :434-435)

plus the geometric grid, exponential porosity and "constant then
exponential" bioturbation profile that the benchmark configurations use.
All arrays are per column, top layer first.  ``dzc[k]`` is the distance
between the centres of layers k-1 and k (the reference's ``dzc(k-1)``);
``dzc[0]`` is unused and set to 0.
"""
from __future__ import annotations

import numpy as np


def init_grid_linear(K: int, dzmin: float):
    """Reference grid: layer heights grow linearly from dzmin, total depth
    0.18 K/(K+1) m (fabm_sediment_driver.F90:147,159-168)."""
    fac = 0.18 / ((K + 1) / 2.0 * dzmin) - 1.0
    dz = np.empty(K)
    for k in range(1, K + 1):
        dz[k - 1] = (1.0 + (fac - 1.0) * (k - 1) / (K - 1)) * dzmin
    return _centres(dz)


def geometric_ratio(K: int, dz1: float, total: float) -> float:
    """Ratio r with dz1 (r^K - 1)/(r - 1) = total (bisection to full
    precision).  A single layer has no ratio (its layer is dz1 whatever the
    total): 1."""
    if K <= 1 or abs(K * dz1 - total) <= 1e-15 * total:
        return 1.0
    def s(r):
        return dz1 * K if r == 1.0 else dz1 * (r ** K - 1.0) / (r - 1.0)
    lo, hi = (1.0, 2.0) if K * dz1 < total else (1e-6, 1.0)
    while s(hi) < total:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if s(mid) < total:
            lo = mid
        else:
            hi = mid
        if hi - lo <= 1e-16 * hi:
            break
    return 0.5 * (lo + hi)


def geometric_ratios(K, dz1: float, total):
    """geometric_ratio for arrays of layer counts and totals: the same
    bisection, element by element (each element stops at the step the
    scalar version stops at), so the ratios are the same numbers."""
    K = np.asarray(K, dtype=np.int64)
    total = np.asarray(total, dtype=np.float64)
    K, total = np.broadcast_arrays(K, total)
    out = np.ones(K.shape)
    uniform = (K <= 1) | (np.abs(K * dz1 - total) <= 1e-15 * total)
    todo = ~uniform
    if not todo.any():
        return out
    Kf, T = K[todo].astype(np.float64), total[todo]
    Kint = K[todo]

    def s(r):
        one = r == 1.0
        v = np.where(one, dz1 * Kf, 0.0)
        rr = np.where(one, 2.0, r)
        return np.where(one, v, dz1 * (rr ** Kint - 1.0) / (rr - 1.0))

    grow = Kint * dz1 < T
    lo = np.where(grow, 1.0, 1e-6)
    hi = np.where(grow, 2.0, 1.0)
    while True:
        short = s(hi) < T
        if not short.any():
            break
        hi = np.where(short, hi * 2.0, hi)
    live = np.ones(lo.shape, dtype=bool)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        below = s(mid) < T
        lo = np.where(live & below, mid, lo)
        hi = np.where(live & ~below, mid, hi)
        live &= ~(hi - lo <= 1e-16 * hi)
        if not live.any():
            break
    out[todo] = 0.5 * (lo + hi)
    return out


def init_grid_geometric(K: int, dz1: float, total: float):
    """Geometric layers: top layer dz1, constant growth ratio, given total."""
    r = geometric_ratio(K, dz1, total)
    dz = dz1 * r ** np.arange(K, dtype=np.float64)
    return _centres(dz)


def _centres(dz: np.ndarray):
    K = dz.size
    zi = np.empty(K + 1)
    zc = np.empty(K)
    zi[0] = 0.0
    for k in range(K):
        zc[k] = zi[k] + 0.5 * dz[k]
        zi[k + 1] = zi[k] + dz[k]
    dzc = np.zeros(K)
    dzc[1:] = zc[1:] - zc[:-1]
    return {"dz": dz, "zc": zc, "zi": zi, "dzc": dzc}


def porosity_linear(zc, porosity_max=0.7, porosity_fac=0.9):
    """fabm_sediment_driver.F90:280."""
    return porosity_max * (1.0 - porosity_fac * np.asarray(zc))


def porosity_exponential(zc, phi0, phi_inf, length=0.02):
    """phi(z) = phi_inf + (phi0 - phi_inf) exp(-z / length)."""
    return phi_inf + (phi0 - phi_inf) * np.exp(-np.asarray(zc) / length)


def interface_porosity(phi):
    """update_porosity :434-435: phiI_1 = phi_1, phiI_k = (phi_{k-1}+phi_k)/2."""
    phi = np.asarray(phi, dtype=np.float64)
    out = np.empty_like(phi)
    out[0] = phi[0]
    out[1:] = 0.5 * (phi[:-1] + phi[1:])
    return out


def bioturbation_factor(zi_upper, profile, bioturbation=0.9, depth_cm=5.0, min_cm2d=0.2):
    """Bioturbation factor at the upper interface of each layer
    (fabm_sediment_driver.F90:293-303): 0 constant, 1 linear decrease to
    bioturbation_min, 2 exponential."""
    zi = np.asarray(zi_upper, dtype=np.float64)
    if profile == 0:
        return np.ones_like(zi)
    if profile == 1:
        return np.maximum(min_cm2d / bioturbation, np.maximum(depth_cm - 100.0 * zi, 0.0) / depth_cm)
    if profile == 2:
        return np.exp(-100.0 * zi / depth_cm)
    raise ValueError("bioturbation profile 3 (Zhang & Wirtz) is not supported")


def bioturbation_const_exp(zi_upper, mixed_depth=0.05, length=0.05):
    """1 down to mixed_depth, then exp(-(z - mixed_depth)/length)."""
    zi = np.asarray(zi_upper, dtype=np.float64)
    return np.where(zi <= mixed_depth, 1.0, np.exp(-(zi - mixed_depth) / length))


def update_porosity(phi_surface, zc, dz, off, porosity_fac=0.9, pom_flux_max=2.0e4, mask=None):
    """update_porosity(from_surface=.true.) of fabm_sediment_driver.F90:
    393-442 on packed columns (column c owns cells off[c]..off[c+1]-1).
    The caller's porosity import sets the top layer (fabm_sediment_component.
    F90:1642-1643); then per column
      phi_k = phi_1 (1 - porosity_fac (zc_k - zc_1)),  k >= 2   (:410-413)
      flux_cap_k = pom_flux_max/86400 (1 - phi_k) dz_k, not increasing
      below the second layer                                    (:415-426)
      phi = 1 on masked columns                                 (:431)
      interface porosity phiI_1 = phi_1, phiI_k = (phi_k-1 + phi_k)/2 (:434-435).
    phi_surface is per column; returns (phi, phiI, flux_cap) per cell, to be
    handed to sedcol_set_geometry / sedcol_set_flux_cap."""
    zc = np.asarray(zc, dtype=np.float64)
    dz = np.asarray(dz, dtype=np.float64)
    off = np.asarray(off, dtype=np.int64)
    ncol = off.size - 1
    phi_surface = np.broadcast_to(np.asarray(phi_surface, dtype=np.float64), (ncol,))
    phi = np.empty_like(zc)
    phiI = np.empty_like(zc)
    fc = np.empty_like(zc)
    for c in range(ncol):
        sl = slice(off[c], off[c + 1])
        z = zc[sl]
        p = np.empty(z.size)
        p[0] = phi_surface[c]
        p[1:] = phi_surface[c] * (1.0 - porosity_fac * (z[1:] - z[0]))
        fc[sl] = flux_cap(p, dz[sl], pom_flux_max)
        if mask is not None and mask[c]:
            p[:] = 1.0
        phi[sl] = p
        phiI[sl] = interface_porosity(p)
    return phi, phiI, fc


def flux_cap(phi, dz, pom_flux_max=2.0e4):
    """Distributed-POM flux cap (fabm_sediment_driver.F90:284-291)."""
    fc = pom_flux_max / 86400.0 * (1.0 - np.asarray(phi)) * np.asarray(dz)
    for k in range(2, fc.size):
        if fc[k] > fc[k - 1]:
            fc[k] = fc[k - 1]
    return fc
