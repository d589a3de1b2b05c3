"""ctypes binding of libsedcol.so (include/sedcol.h) and the host-side
mirror of the reference driver interface.

``SedimentColumns`` plays the role of ``type_sed`` (src/drivers/
fabm_sediment_driver.F90:69-113) together with ``ode_solver``
(src/utilities/solver_library.F90:80-189): the same set-up calls, the same
``get_rhs`` / ``ode_solver`` semantics, with the arrays living on the GPU.
Errors raise ``SedcolError`` carrying the library's status code instead of
the reference's ``stop``.

The library is loaded from this package directory; there is no fallback:
a missing or unloadable libsedcol.so raises at import of this module's
``lib()``.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import NamedTuple

import numpy as np

from .params import PARAM_FIELDS, NVAR, SedParams, fill_struct, params_dict

LIB_NAME = "libsedcol.so"
HERE = Path(__file__).resolve().parent

SEDCOL_OK, SEDCOL_EINVAL, SEDCOL_EDOMAIN, SEDCOL_ENAN = 0, 1, 2, 3
SEDCOL_EHIP, SEDCOL_ELIMIT, SEDCOL_ENOMEM, SEDCOL_EUNSUPPORTED = 4, 5, 6, 7
LAYOUT_FORTRAN, LAYOUT_PACKED = 0, 1
NBUD = 16

#: every symbol include/sedcol.h declares
EXPORTS = (
    "sedcol_create", "sedcol_destroy", "sedcol_last_error", "sedcol_set_params", "sedcol_set_geometry",
    "sedcol_set_forcing", "sedcol_set_forcing_series", "sedcol_set_state", "sedcol_get_state",
    "sedcol_set_state_profiles", "sedcol_rhs",
    "sedcol_rhs_noalloc", "sedcol_rates",
    "sedcol_step", "sedcol_run", "sedcol_integrate", "sedcol_get_fluxes", "sedcol_get_budget",
    "sedcol_reset_budget", "sedcol_get_inventory", "sedcol_get_diagnostics", "sedcol_reset_diagnostics",
    "sedcol_residual", "sedcol_get_last_min_dt", "sedcol_synchronize", "sedcol_last_timing",
    "sedcol_set_schedule", "sedcol_get_schedule", "sedcol_set_tuning", "sedcol_get_tuning", "sedcol_autotune",
    "sedcol_checkpoint_size", "sedcol_save_checkpoint", "sedcol_load_checkpoint", "sedcol_spinup",
    "sedcol_assemble_forcing", "sedcol_set_flux_cap", "sedcol_set_bioturbation_zw",
)
NBIN = 3


class DeviceRows(NamedTuple):
    """rows x ld float32 values in device memory (row-major), owned elsewhere."""
    ptr: int
    rows: int
    ld: int


class SedcolError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sedcol status {code}: {msg}")
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [("ncol", ctypes.c_int32), ("kmax", ctypes.c_int32), ("nvar", ctypes.c_int32),
                ("layout", ctypes.c_int32), ("nlayers", ctypes.POINTER(ctypes.c_int32)),
                ("mask", ctypes.POINTER(ctypes.c_int32))]


class Params(ctypes.Structure):
    _fields_ = PARAM_FIELDS


class ColDiag(ctypes.Structure):
    _fields_ = [("accepted", ctypes.c_int64), ("rejected", ctypes.c_int64), ("forced", ctypes.c_int64),
                ("zero_sink", ctypes.c_int64), ("newton", ctypes.c_int64), ("dt_last", ctypes.c_double), ("dt_next", ctypes.c_double),
                ("dt_minacc", ctypes.c_double), ("p_min", ctypes.c_double), ("status", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


class Schedule(ctypes.Structure):
    _fields_ = [("window", ctypes.c_int32), ("graph_windows", ctypes.c_int32), ("check_every", ctypes.c_int32),
                ("concurrent_bins", ctypes.c_int32)]


class BioturbationZw(ctypes.Structure):
    _fields_ = [("k_l", ctypes.c_double), ("L1", ctypes.c_double), ("L2", ctypes.c_double),
                ("beta", ctypes.c_double), ("b", ctypes.c_double), ("dry_density", ctypes.c_double)]


#: &sed_nml defaults of bioturbation profile 3 (fabm_sediment_driver.F90:226-231):
#: k_l, L1, L2, beta, b, dry density
ZW_DEFAULTS = (0.11, 0.2, 0.6, 0.22, 1.334, 1000.0)


class Tuning(ctypes.Structure):
    _fields_ = [("lanes_per_column", ctypes.c_int32 * 3), ("columns_per_group", ctypes.c_int32 * 3),
                ("layers_per_lane", ctypes.c_int32 * 3)]

    def as_dict(self) -> dict:
        return {"lanes_per_column": list(self.lanes_per_column), "columns_per_group": list(self.columns_per_group),
                "layers_per_lane": list(self.layers_per_lane)}


DIAG_DTYPE = np.dtype([("accepted", np.int64), ("rejected", np.int64), ("forced", np.int64),
                       ("zero_sink", np.int64), ("newton", np.int64), ("dt_last", np.float64), ("dt_next", np.float64),
                       ("dt_minacc", np.float64), ("p_min", np.float64), ("status", np.int32),
                       ("pad", np.int32)])

_LIB = None


def lib_path() -> Path:
    return Path(os.environ.get("SEDCOL_LIB", HERE / LIB_NAME))


def lib() -> ctypes.CDLL:
    """Load libsedcol.so (raises OSError if it is missing)."""
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not path.exists():
            raise OSError(f"{path} not found: build it with __graft_entry__.build()")
        L = ctypes.CDLL(str(path))
        vp, i32, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
        pd = ctypes.POINTER(ctypes.c_double)
        sig = {
            "sedcol_create": (ctypes.c_int, [ctypes.POINTER(Desc), ctypes.POINTER(vp)]),
            "sedcol_destroy": (ctypes.c_int, [vp]),
            "sedcol_last_error": (ctypes.c_char_p, []),
            "sedcol_set_params": (ctypes.c_int, [vp, ctypes.POINTER(Params)]),
            "sedcol_set_geometry": (ctypes.c_int, [vp, pd, pd, pd, pd, pd, pd]),
            "sedcol_set_forcing": (ctypes.c_int, [vp, pd, pd, pd]),
            "sedcol_set_forcing_series": (ctypes.c_int, [vp, i32, dbl, dbl, pd, pd, pd, i32]),
            "sedcol_set_state": (ctypes.c_int, [vp, pd]),
            "sedcol_get_state": (ctypes.c_int, [vp, pd]),
            "sedcol_set_state_profiles": (ctypes.c_int, [vp, vp, ctypes.c_int64, i32, pd, pd, pd, dbl, pd]),
            "sedcol_rhs": (ctypes.c_int, [vp, pd, pd, pd]),
            "sedcol_rhs_noalloc": (ctypes.c_int, [vp, pd, pd, pd]),
            "sedcol_rates": (ctypes.c_int, [vp, pd, pd]),
            "sedcol_step": (ctypes.c_int, [vp, dbl, i32, pd, pd]),
            "sedcol_run": (ctypes.c_int, [vp, dbl, i32, i32, i32]),
            "sedcol_integrate": (ctypes.c_int, [vp, dbl, dbl]),
            "sedcol_get_fluxes": (ctypes.c_int, [vp, pd]),
            "sedcol_get_budget": (ctypes.c_int, [vp, pd]),
            "sedcol_reset_budget": (ctypes.c_int, [vp]),
            "sedcol_get_inventory": (ctypes.c_int, [vp, pd]),
            "sedcol_get_diagnostics": (ctypes.c_int, [vp, vp]),
            "sedcol_reset_diagnostics": (ctypes.c_int, [vp]),
            "sedcol_residual": (ctypes.c_int, [vp, pd, pd]),
            "sedcol_get_last_min_dt": (ctypes.c_int, [vp, pd]),
            "sedcol_synchronize": (ctypes.c_int, [vp]),
            "sedcol_last_timing": (ctypes.c_int, [vp, pd, ctypes.POINTER(ctypes.c_int32)]),
            "sedcol_set_schedule": (ctypes.c_int, [vp, ctypes.POINTER(Schedule)]),
            "sedcol_get_schedule": (ctypes.c_int, [vp, ctypes.POINTER(Schedule)]),
            "sedcol_set_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
            "sedcol_get_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
            "sedcol_autotune": (ctypes.c_int, [vp, dbl, dbl, i32, ctypes.POINTER(Tuning), pd]),
            "sedcol_spinup": (ctypes.c_int, [vp, dbl, dbl, dbl, dbl, pd, pd, dbl, pd, pd, pd,
                                             ctypes.POINTER(ctypes.c_int32)]),
            "sedcol_checkpoint_size": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int64)]),
            "sedcol_assemble_forcing": (ctypes.c_int, [vp, pd, pd, pd, ctypes.POINTER(ctypes.c_int32), pd]),
            "sedcol_set_flux_cap": (ctypes.c_int, [vp, pd]),
            "sedcol_set_bioturbation_zw": (ctypes.c_int, [vp, ctypes.POINTER(BioturbationZw)]),
            "sedcol_save_checkpoint": (ctypes.c_int, [vp, vp, ctypes.c_int64]),
            "sedcol_load_checkpoint": (ctypes.c_int, [vp, vp, ctypes.c_int64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _pd(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def check(code: int):
    if code != SEDCOL_OK:
        msg = lib().sedcol_last_error()
        raise SedcolError(code, msg.decode() if msg else "")


class SedimentColumns:
    """GPU sediment domain with the reference driver's call sequence.

    layout "packed": arrays as in ``cases.Domain`` (state (NV, ncell));
    layout "fortran": uniform K, state (NV, K, ncol) = Fortran conc(i,j,k,n)
    with ncol = I*J, cell arrays (K, ncol).
    """

    def __init__(self, nlayers, mask=None, layout="packed"):
        self.nlayers = _c(nlayers, np.int32)
        self.ncol = int(self.nlayers.size)
        self.kmax = int(self.nlayers.max()) if self.nlayers.size else 0
        # an unknown layout name goes to the library as -1, which rejects it
        self.layout = {"packed": LAYOUT_PACKED, "fortran": LAYOUT_FORTRAN}.get(layout, -1)
        self.mask = _c(np.zeros(self.ncol) if mask is None else mask, np.int32)
        self.ncell = int(self.nlayers.sum())
        d = Desc()
        d.ncol, d.kmax, d.nvar, d.layout = self.ncol, self.kmax, NVAR, self.layout
        # the Fortran layout takes its uniform K from kmax; ragged counts are
        # passed on so that the library rejects them
        uniform = self.nlayers.size > 0 and np.all(self.nlayers == self.kmax)
        d.nlayers = (None if self.layout == LAYOUT_FORTRAN and uniform
                     else self.nlayers.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        d.mask = self.mask.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        h = ctypes.c_void_p()
        check(lib().sedcol_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h

    # -- lifecycle -----------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().sedcol_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- set-up --------------------------------------------------------
    def set_params(self, p: SedParams | dict):
        s = fill_struct(Params, params_dict(p))
        check(lib().sedcol_set_params(self._h, ctypes.byref(s)))

    def set_geometry(self, dz, dzc, phi, phiI, bfac, biot):
        arrs = [_c(a) for a in (dz, dzc, phi, phiI, bfac, biot)]
        check(lib().sedcol_set_geometry(self._h, *[_pd(a) for a in arrs]))

    def set_forcing(self, temp, cbw, fup):
        t, c, f = _c(temp), _c(cbw), _c(fup)
        check(lib().sedcol_set_forcing(self._h, _pd(t), _pd(c), _pd(f)))

    def set_forcing_series(self, temp, cbw, fup, t0, dt_rec):
        """temp (nrec, ncol), cbw (nrec, NV, ncol), fup (nrec or 1, NV, ncol)."""
        t, c, f = _c(temp), _c(cbw), _c(fup)
        nrec = int(t.shape[0])
        nrf = int(f.shape[0]) if f.ndim == 3 else 1
        check(lib().sedcol_set_forcing_series(self._h, nrec, float(t0), float(dt_rec), _pd(t), _pd(c), _pd(f), nrf))

    def set_bioturbation_zw(self, zw=ZW_DEFAULTS):
        """Bioturbation profile 3 (Zhang & Wirtz) with (k_l, L1, L2, beta, b,
        dry density); None returns to the set_geometry factor."""
        if zw is None:
            check(lib().sedcol_set_bioturbation_zw(self._h, None))
        else:
            z = BioturbationZw(*(float(v) for v in zw))
            check(lib().sedcol_set_bioturbation_zw(self._h, ctypes.byref(z)))

    def set_flux_cap(self, fcap):
        """Distributed-POM flux cap per cell (bcup_part = 4)."""
        check(lib().sedcol_set_flux_cap(self._h, _pd(_c(fcap))))

    def assemble_forcing(self, present, temp_surface=None, c_surface=None, w_surface=None):
        """Boundary conditions from the fields at the soil surface
        (get_boundary_conditions); returns the top fluxes (NV, ncol)."""
        pr = np.ascontiguousarray(np.asarray(present, dtype=np.int32).reshape(NVAR))
        ts = None if temp_surface is None else _c(temp_surface).reshape(self.ncol)
        cs = None if c_surface is None else _c(c_surface).reshape(NVAR, self.ncol)
        ws = None if w_surface is None else _c(w_surface).reshape(NVAR, self.ncol)
        f = np.empty((NVAR, self.ncol))
        check(lib().sedcol_assemble_forcing(self._h, _pd(ts) if ts is not None else None,
                                            _pd(cs) if cs is not None else None,
                                            _pd(ws) if ws is not None else None,
                                            pr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _pd(f)))
        return f

    def set_state(self, conc):
        x = _c(conc)
        check(lib().sedcol_set_state(self._h, _pd(x)))

    def get_state(self) -> np.ndarray:
        x = np.empty(self._state_shape())
        check(lib().sedcol_get_state(self._h, _pd(x)))
        return x

    def set_state_profiles(self, prof: DeviceRows, nz, scale, offset, eps, dz1=1.0e-3, depth=None):
        """Warm start on the device from log10 profiles (sedcol_set_state_profiles).
        prof: DeviceRows of ncol float32 rows in device memory (for example
        from surrogate.Network.predict_device); scale/offset (NV*nz), eps
        (NV) and depth (ncol, optional) are host arrays."""
        if prof.rows != self.ncol or prof.ld < NVAR * nz:
            raise ValueError("prof must hold ncol rows of at least nvar*nz values")
        sc, of = _c(scale), _c(offset)
        ep = _c(eps)
        if sc.size != NVAR * nz or of.size != NVAR * nz or ep.size != NVAR:
            raise ValueError("scale and offset need nvar*nz entries, eps nvar")
        dp = None if depth is None else _c(depth)
        if dp is not None and dp.size != self.ncol:
            raise ValueError("depth needs ncol entries")
        check(lib().sedcol_set_state_profiles(self._h, ctypes.c_void_p(prof.ptr), int(prof.ld),
                                              int(nz), _pd(sc), _pd(of), _pd(ep), float(dz1),
                                              _pd(dp) if dp is not None else None))

    def _state_shape(self):
        if self.layout == LAYOUT_PACKED:
            return (NVAR, self.ncell)
        return (NVAR, self.kmax, self.ncol)

    # -- hot path ------------------------------------------------------
    def get_rhs(self, conc):
        """get_rhs: (rhs, top fluxes (NV, ncol))."""
        x = _c(conc)
        r = np.empty_like(x)
        f = np.empty((NVAR, self.ncol))
        check(lib().sedcol_rhs(self._h, _pd(x), _pd(r), _pd(f)))
        return r, f

    def rates(self):
        """(DIC release, N2-N release) per cell of the resident state
        (mmol m-3 pore water s-1), cell layout of the handle."""
        shape = (self.ncell,) if self.layout == LAYOUT_PACKED else (self.kmax, self.ncol)
        a = np.empty(shape)
        b = np.empty(shape)
        check(lib().sedcol_rates(self._h, _pd(a), _pd(b)))
        return a, b

    def get_rhs_noalloc(self, conc):
        """get_rhs through sedcol_rhs_noalloc (no allocation inside the call)."""
        x = _c(conc)
        r = np.empty_like(x)
        f = np.empty((NVAR, self.ncol))
        check(lib().sedcol_rhs_noalloc(self._h, _pd(x), _pd(r), _pd(f)))
        return r, f

    def ode_solver(self, conc, dt, method):
        """One ode_solver call on a host state, updated in place; returns
        the top fluxes of the last right-hand side."""
        x = conc if (conc.flags.c_contiguous and conc.dtype == np.float64) else _c(conc)
        f = np.empty((NVAR, self.ncol))
        check(lib().sedcol_step(self._h, float(dt), int(method), _pd(x), _pd(f)))
        if x is not conc:
            conc[...] = x
        return f

    def run(self, dt, nsteps, method, steps_per_launch=0):
        check(lib().sedcol_run(self._h, float(dt), int(nsteps), int(method), int(steps_per_launch)))

    def integrate(self, t_start, t_end):
        check(lib().sedcol_integrate(self._h, float(t_start), float(t_end)))

    # -- schedule and kernel shape -------------------------------------
    def set_schedule(self, window=0, graph_windows=0, check_every=1, concurrent_bins=0):
        s = Schedule(int(window), int(graph_windows), int(check_every), int(concurrent_bins))
        check(lib().sedcol_set_schedule(self._h, ctypes.byref(s)))

    def get_schedule(self) -> dict:
        s = Schedule()
        check(lib().sedcol_get_schedule(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in Schedule._fields_}

    def set_tuning(self, lanes_per_column, columns_per_group, layers_per_lane=(1, 1, 1)):
        t = Tuning((ctypes.c_int32 * 3)(*lanes_per_column), (ctypes.c_int32 * 3)(*columns_per_group),
                   (ctypes.c_int32 * 3)(*layers_per_lane))
        check(lib().sedcol_set_tuning(self._h, ctypes.byref(t)))

    def get_tuning(self) -> dict:
        t = Tuning()
        check(lib().sedcol_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def autotune(self, t_start, t_span, repeats=2):
        """Time every kernel shape per layer bin, install the fastest.

        Returns (tuning dict, ms table of shape (3 bins, 2 layers per lane
        1/2, 3 lanes 16/32/64, 3 workgroups 64/128/256), -1 where a shape
        does not apply)."""
        t = Tuning()
        table = np.empty((NBIN, 2, 3, 3))
        check(lib().sedcol_autotune(self._h, float(t_start), float(t_span), int(repeats), ctypes.byref(t),
                                    _pd(table)))
        return t.as_dict(), table

    def spinup(self, t_start, chunk, t_max, tol, cfloor, flux_ref=None, flux_rtol=0.01, flux_floor=None):
        """Per-column adaptive-EMP spin-up; returns (t_conv, resid,
        n_unconverged): time after t_start at which each column converged
        (residual below tol and, with flux_ref (NV, ncol), top fluxes within
        flux_rtol * max(|flux_ref|, flux_floor)), -1 if never / land; final
        residual per column."""
        cf = _c(cfloor)
        tc = np.empty(self.ncol)
        rs = np.empty(self.ncol)
        nu = ctypes.c_int32()
        fr = ff = None
        if flux_ref is not None:
            fr = _c(flux_ref).reshape(NVAR, self.ncol)
            ff = _c(np.broadcast_to(0.0 if flux_floor is None else flux_floor, (self.ncol,)))
        check(lib().sedcol_spinup(self._h, float(t_start), float(chunk), float(t_max), float(tol), _pd(cf),
                                  _pd(fr) if fr is not None else None, float(flux_rtol),
                                  _pd(ff) if ff is not None else None, _pd(tc), _pd(rs), ctypes.byref(nu)))
        return tc, rs, nu.value

    # -- restart -------------------------------------------------------
    def save_checkpoint(self) -> bytes:
        n = ctypes.c_int64()
        check(lib().sedcol_checkpoint_size(self._h, ctypes.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        check(lib().sedcol_save_checkpoint(self._h, ctypes.c_void_p(buf.ctypes.data), n.value))
        return buf.tobytes()

    def load_checkpoint(self, blob: bytes):
        buf = np.frombuffer(blob, dtype=np.uint8)
        check(lib().sedcol_load_checkpoint(self._h, ctypes.c_void_p(buf.ctypes.data), buf.size))

    # -- outputs -------------------------------------------------------
    def fluxes(self):
        f = np.empty((NVAR, self.ncol))
        check(lib().sedcol_get_fluxes(self._h, _pd(f)))
        return f

    def budget(self):
        b = np.empty((self.ncol, NBUD))
        check(lib().sedcol_get_budget(self._h, _pd(b)))
        return b

    def reset_budget(self):
        check(lib().sedcol_reset_budget(self._h))

    def inventory(self):
        v = np.empty((NVAR, self.ncol))
        check(lib().sedcol_get_inventory(self._h, _pd(v)))
        return v

    def diagnostics(self):
        d = np.zeros(self.ncol, dtype=DIAG_DTYPE)
        check(lib().sedcol_get_diagnostics(self._h, ctypes.c_void_p(d.ctypes.data)))
        return d

    def reset_diagnostics(self):
        check(lib().sedcol_reset_diagnostics(self._h))

    def residual(self, cfloor):
        cf = _c(cfloor)
        r = np.empty(self.ncol)
        check(lib().sedcol_residual(self._h, _pd(cf), _pd(r)))
        return r

    def last_min_dt(self):
        v = ctypes.c_double()
        check(lib().sedcol_get_last_min_dt(self._h, ctypes.byref(v)))
        return v.value

    def last_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_int32()
        check(lib().sedcol_last_timing(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # -- convenience ---------------------------------------------------
    @classmethod
    def from_domain(cls, dom, record=0, series=None):
        """Create, configure and load a cases.Domain (packed layout).  With
        more than one forcing record the whole series is set, unless series
        is False: then record `record` is the current forcing."""
        s = cls(dom.nlayers, dom.mask, layout="packed")
        s.set_params(dom.params)
        s.set_geometry(dom.dz, dom.dzc, dom.phi, dom.phiI, dom.bfac, dom.biot)
        if dom.nrec > 1 and series is not False:
            s.set_forcing_series(dom.temp, dom.cbw, dom.fup, dom.t0, dom.dt_rec)
        else:
            s.set_forcing(*dom.record(record))
        s.set_state(dom.conc0)
        return s
