"""sedcol -- MI355X sediment-column hot path (host side).

The compute lives in libsedcol.so (HIP, gfx950) behind the C ABI of
include/sedcol.h; this package builds it, binds it with ctypes, and mirrors
the reference driver interface (type_sed / ode_solver) for Python callers.
"""
from .params import (VARIABLES, NVAR, PARTICULATE, ELEMENT_VARS, SedParams, OmexdiaNamelist,
                     EULER, RUNGE_KUTTA_4, ADAPTIVE_EULER, RUNGE_KUTTA_4_38, EMP1, EMP2, EMP_ADAPTIVE)

__all__ = [
    "VARIABLES", "NVAR", "PARTICULATE", "ELEMENT_VARS", "SedParams", "OmexdiaNamelist",
    "EULER", "RUNGE_KUTTA_4", "ADAPTIVE_EULER", "RUNGE_KUTTA_4_38", "EMP1", "EMP2", "EMP_ADAPTIVE",
]
