"""File formats around the sediment-column library: the standalone
driver's text output and the `*_in_soil` restart fields.

* ``output.dat`` (examples/standalone/omexdia_p/main.F90:109-157): a header
  line ``time(s) depth(m) layer-height(m) porosity()`` followed by the state
  variable and diagnostic names; then, every output interval, one line
  ``<time> fluxes <f_1> ... <f_nvar>`` (list-directed in the reference, so
  only whitespace-separated numbers are promised) and one line per layer in
  the fixed format ``(E15.3, A, E15.4E3, A, E15.4E3, A, E15.4E3)`` for time,
  layer centre depth, layer height and porosity, then ``(A, E15.4E3)`` per
  state and diagnostic variable.  The reference's reader
  (examples/standalone/omexdia_p/plotsed1d.py) splits lines on whitespace,
  recognises flux lines by their second token, and finds porosity and the
  denitrification diagnostic by their header position.
* restart (fabm_sediment_component.F90:1328-1489): for each state variable
  the import state may hold a field ``<name>_in_soil`` of the state's
  (I, J, K) shape; a complete field of matching bounds replaces that
  variable's concentrations (:1462-1463), anything else is skipped and the
  state kept.  The export side (:1776-1819) writes the same names.  Fields
  are stored as NetCDF-3 classic files (scipy.io.netcdf_file).

Names are the omexdia_p variable names the reference drivers recognise
(main.F90:90-101, fabm_sediment_component.F90:589-604) in library order; the
adsorbed phosphorus pool has no reference name and is called
``phosphate_adsorbed``.
"""
from __future__ import annotations

import math

import numpy as np

from .params import NVAR

#: state variable names in library order
STATE_NAMES = (
    "detritus_labile_carbon", "detritus_semilabile_carbon", "detritus_labile_nitrogen",
    "detritus_semilabile_nitrogen", "detritus_labile_phosphorus", "dissolved_oxygen", "dissolved_nitrate",
    "dissolved_ammonium", "dissolved_reduced_substances", "dissolved_phosphate", "phosphate_adsorbed",
)
#: the denitrification diagnostic the reference's plotting script reads
DENIT_NAME = "hzg_omexdia_p_denit"
assert len(STATE_NAMES) == NVAR


# ---------------------------------------------------------------------------
# Fortran E edit descriptors
# ---------------------------------------------------------------------------

def fortran_e(x: float, w: int, d: int, e: int | None = None) -> str:
    """Fortran ``Ew.d`` (``e`` None) or ``Ew.dEe`` output of x: a 0.d...d
    mantissa with d digits, the exponent with e digits (Ew.d: 2 digits and
    the letter E up to 99, else 3 digits without it), right-justified in w
    characters; w asterisks if it does not fit."""
    x = float(x)
    if math.isnan(x) or math.isinf(x):
        s = "NaN" if math.isnan(x) else ("-Infinity" if x < 0 else "Infinity")
        return s.rjust(w) if len(s) <= w else "*" * w
    if x == 0.0:
        mant, ex = "0." + "0" * d, 0
        sign = "-" if math.copysign(1.0, x) < 0 else ""
    else:
        sign = "-" if x < 0 else ""
        a = abs(x)
        # decimal digits by exact rounding of the value to d significant digits
        s = f"{a:.{d - 1}e}"          # d.ddd e+XX, correctly rounded
        m, ex = s.split("e")
        digits = m.replace(".", "")
        ex = int(ex) + 1              # 0.dddd scaling
        mant = "0." + digits
    if e is None:
        if abs(ex) <= 99:
            expo = f"E{'-' if ex < 0 else '+'}{abs(ex):02d}"
        elif abs(ex) <= 999:
            expo = f"{'-' if ex < 0 else '+'}{abs(ex):03d}"
        else:
            return "*" * w
    else:
        if abs(ex) >= 10 ** e:
            return "*" * w
        expo = f"E{'-' if ex < 0 else '+'}{abs(ex):0{e}d}"
    out = sign + mant + expo
    return out.rjust(w) if len(out) <= w else "*" * w


# ---------------------------------------------------------------------------
# output.dat
# ---------------------------------------------------------------------------

class OutputDat:
    """Writer of the standalone driver's ``output.dat`` for one column
    (main.F90:109-157)."""

    def __init__(self, path, state_names=STATE_NAMES, diagnostic_names=(DENIT_NAME,)):
        self.f = open(path, "w")
        self.ndiag = len(diagnostic_names)
        head = "time(s) " + "depth(m) " + "layer-height(m) " + "porosity() "
        head += "".join(" " + n for n in state_names)
        head += "".join(" " + n for n in diagnostic_names)
        self.f.write(head + "\n")

    def write(self, t, fluxes, zc, dz, porosity, conc, diagnostics=()):
        """t (s); fluxes (nvar) top fluxes; zc, dz, porosity (K); conc
        (nvar, K); diagnostics (ndiag, K)."""
        self.f.write(" " + repr(float(t)) + " fluxes" + "".join(" " + repr(float(v)) for v in fluxes) + "\n")
        conc = np.asarray(conc)
        diag = np.asarray(diagnostics) if self.ndiag else np.zeros((0, conc.shape[1]))
        for k in range(conc.shape[1]):
            line = (fortran_e(t, 15, 3) + " " + fortran_e(zc[k], 15, 4, 3) + " " + fortran_e(dz[k], 15, 4, 3)
                    + " " + fortran_e(porosity[k], 15, 4, 3))
            line += "".join(" " + fortran_e(v, 15, 4, 3) for v in conc[:, k])
            line += "".join(" " + fortran_e(v, 15, 4, 3) for v in diag[:, k])
            self.f.write(line + "\n")

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def denit_diagnostic(n2_release):
    """hzg_omexdia_p_denit of the output (mmol N m-3 d-1) from the library's
    N2-N release rate (mmol N m-3 s-1, sedcol_rates)."""
    return np.asarray(n2_release) * 86400.0


# ---------------------------------------------------------------------------
# export fields
# ---------------------------------------------------------------------------

def upward_flux_fields(flux_top, ni: int, nj: int, names=STATE_NAMES) -> dict:
    """Export side of the top fluxes (fabm_sediment_component.F90:1776-1819):
    {<name>_upward_flux_at_soil_surface: (I, J) array} from the library's
    top fluxes (nvar, ncol), column c = i + I j.  The library's fluxes are
    positive downward, the export positive upward: the field is -flux
    (:1798, 1812)."""
    f = np.asarray(flux_top, dtype=np.float64)
    if f.shape != (len(names), ni * nj):
        raise ValueError(f"flux_top must be ({len(names)}, {ni * nj})")
    return {f"{n}_upward_flux_at_soil_surface": np.ascontiguousarray((0.0 - f[v]).reshape(nj, ni).T)
            for v, n in enumerate(names)}


# ---------------------------------------------------------------------------
# restart fields
# ---------------------------------------------------------------------------

def restart_fields(conc_ijkn: np.ndarray, names=STATE_NAMES) -> dict:
    """Export side: {name_in_soil: conc(:, :, :, n)} of a state (I, J, K, nvar)
    (fabm_sediment_component.F90:1776-1819)."""
    conc_ijkn = np.asarray(conc_ijkn, dtype=np.float64)
    return {f"{n}_in_soil": np.ascontiguousarray(conc_ijkn[..., i]) for i, n in enumerate(names)}


def apply_restart_fields(fields: dict, conc_ijkn: np.ndarray, names=STATE_NAMES):
    """Import side (fabm_sediment_component.F90:1362-1463): every variable
    whose ``<name>_in_soil`` field is present, complete (finite) and of the
    state's (I, J, K) bounds is replaced; the others keep their values.
    Returns the new state and {field name: 'restored' | reason skipped}."""
    out = np.array(conc_ijkn, dtype=np.float64, copy=True)
    report = {}
    for i, n in enumerate(names):
        key = f"{n}_in_soil"
        if key not in fields:
            report[key] = "not found"
            continue
        a = np.asarray(fields[key], dtype=np.float64)
        if a.shape != out.shape[:3]:
            report[key] = f"skipped: bounds {a.shape} do not match {out.shape[:3]}"
            continue
        if not np.all(np.isfinite(a)):
            report[key] = "skipped: incomplete field"
            continue
        out[..., i] = a
        report[key] = "restored"
    return out, report


def save_restart(path, fields: dict, time_s: float | None = None):
    """NetCDF-3 file with one (k, j, i) variable per field, C order of the
    Fortran (i, j, k) array (the layout of the reference's NetCDF output)."""
    from scipy.io import netcdf_file

    shapes = {np.asarray(v).shape for v in fields.values()}
    if len(shapes) != 1:
        raise ValueError("restart fields must share one (I, J, K) shape")
    I, J, K = shapes.pop()
    with netcdf_file(str(path), "w") as nc:
        nc.createDimension("ungridded00024", K)
        nc.createDimension("y", J)
        nc.createDimension("x", I)
        if time_s is not None:
            nc.createDimension("time", 1)
            tv = nc.createVariable("time", "d", ("time",))
            tv[:] = [time_s]
            tv.units = "s"
        for name, arr in fields.items():
            v = nc.createVariable(name, "d", ("ungridded00024", "y", "x"))
            v[:] = np.asarray(arr, dtype=np.float64).transpose(2, 1, 0)


def load_restart(path) -> dict:
    """Fields of a file written by save_restart, as (I, J, K) arrays."""
    from scipy.io import netcdf_file

    out = {}
    with netcdf_file(str(path), "r", mmap=False) as nc:
        for name, v in nc.variables.items():
            if name.endswith("_in_soil") and len(v.shape) == 3:
                out[name] = np.array(v[:], dtype=np.float64).transpose(2, 1, 0)
    return out
