"""Parameter sets of the sediment-column model.

Transport and solver keys follow the reference namelists:
``&sed_nml`` (src/drivers/fabm_sediment_driver.F90:211-233, defaults
:215-231), ``&run_nml`` solver settings (solver_library.F90:40-41,
examples/standalone/omexdia_p/main.F90:42-50).  Kinetic parameters follow
``&hzg_omexdia_p`` (examples/standalone/omexdia_p/fabm_sed.nml:51-77); rates
given there per day are converted to SI (per second) here, once, so host and
device use identical values.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field, asdict

SECONDS_PER_DAY = 86400.0

#: state variables, in library order (include/sedcol.h enum sedcol_var)
VARIABLES = (
    "ldetC", "sdetC", "ldetN", "sdetN", "detP",
    "O2", "NO3", "NH4", "ODU", "PO4", "adsP",
)
NVAR = len(VARIABLES)
#: particulate variables (solid phase): flux boundary at the top, mixed by
#: bioturbation only (fabm_sediment_driver.F90:658-678)
PARTICULATE = (True, True, True, True, True, False, False, False, False, False, True)
#: element membership used by the budgets
ELEMENT_VARS = {
    "C": (0, 1),
    "N": (2, 3, 6, 7),
    "P": (4, 9, 10),
}

# method ids (solver_library.F90:32-35, plus the EMP methods)
EULER, RUNGE_KUTTA_4, ADAPTIVE_EULER, RUNGE_KUTTA_4_38 = 0, 1, 2, 3
EMP1, EMP2, EMP_ADAPTIVE, ROS2_ADAPTIVE = 4, 5, 6, 7


@dataclass
class OmexdiaNamelist:
    """&hzg_omexdia_p values (per day where the namelist uses days)."""
    rLabile: float = 0.043        # d-1
    rSemilabile: float = 0.001    # d-1
    CprodMax: float = 9600.0      # mmol C m-3 d-1
    rnit: float = 200.0           # d-1
    ksO2nitri: float = 20.0
    rODUox: float = 20.0          # d-1
    ksO2oduox: float = 1.0
    ksO2oxic: float = 3.0
    ksNO3denit: float = 1.0
    kinO2denit: float = 70.0
    kinNO3anox: float = 1.0
    kinO2anox: float = 1.0
    PAds: float = 0.01            # adsorption coefficient (-)
    PAdsODU: float = 70.0         # release half-saturation in ODU (mmol m-3)


@dataclass
class SedParams:
    """All parameters of one sediment domain (mirrors sedcol_params)."""
    D0: float = 0.9                    # cm2 d-1, &sed_nml diffusivity
    dt_min: float = 1.0e-9             # solver_library.F90:40
    relative_change_min: float = -0.9  # solver_library.F90:41
    kin: OmexdiaNamelist = field(default_factory=OmexdiaNamelist)
    emp_rtol: float = 1.0e-6
    emp_atol: float = 1.0e-6
    emp_dt_min: float = 1.0
    emp_dt_max: float = 3600.0
    bcup_dis: int = 2                  # bcup_dissolved_variables
    bcup_part: int = 1                 # bcup_particulate_variables
    bcdown: int = 3
    clamp: int = 1
    arith: int = 0                     # 0 reference operation order, 1 fast form
    emp_weights: int = 0               # Patankar factor: 0 product, 1 geometric mean
    adaptive_method: int = 0           # stepper of integrate/spinup: 0 or 6 EMP, 7 ROS2

    def si(self) -> dict:
        """Flat dict in the field order of sedcol_params / oc_params."""
        k = self.kin
        d = SECONDS_PER_DAY
        return {
            "D0": float(self.D0),
            "dt_min": float(self.dt_min),
            "relative_change_min": float(self.relative_change_min),
            "rL": k.rLabile / d,
            "rS": k.rSemilabile / d,
            "CprodMax": k.CprodMax / d,
            "rnit": k.rnit / d,
            "ksO2nitri": float(k.ksO2nitri),
            "rODUox": k.rODUox / d,
            "ksO2oduox": float(k.ksO2oduox),
            "ksO2oxic": float(k.ksO2oxic),
            "ksNO3denit": float(k.ksNO3denit),
            "kinO2denit": float(k.kinO2denit),
            "kinNO3anox": float(k.kinNO3anox),
            "kinO2anox": float(k.kinO2anox),
            # adsorption onto iron oxides formed by ODU re-oxidation: rate
            # constant PAds * rODUox (library definition)
            "rPads": (k.PAds * k.rODUox) / d,
            "PAdsODU": float(k.PAdsODU),
            "emp_rtol": float(self.emp_rtol),
            "emp_atol": float(self.emp_atol),
            "emp_dt_min": float(self.emp_dt_min),
            "emp_dt_max": float(self.emp_dt_max),
            "bcup_dis": int(self.bcup_dis),
            "bcup_part": int(self.bcup_part),
            "bcdown": int(self.bcdown),
            "clamp": int(self.clamp),
            "arith": int(self.arith),
            "emp_weights": int(self.emp_weights),
            "adaptive_method": int(self.adaptive_method),
        }


PARAM_FIELDS = [
    ("D0", ctypes.c_double), ("dt_min", ctypes.c_double), ("relative_change_min", ctypes.c_double),
    ("rL", ctypes.c_double), ("rS", ctypes.c_double), ("CprodMax", ctypes.c_double),
    ("rnit", ctypes.c_double), ("ksO2nitri", ctypes.c_double), ("rODUox", ctypes.c_double),
    ("ksO2oduox", ctypes.c_double), ("ksO2oxic", ctypes.c_double), ("ksNO3denit", ctypes.c_double),
    ("kinO2denit", ctypes.c_double), ("kinNO3anox", ctypes.c_double), ("kinO2anox", ctypes.c_double),
    ("rPads", ctypes.c_double), ("PAdsODU", ctypes.c_double),
    ("emp_rtol", ctypes.c_double), ("emp_atol", ctypes.c_double),
    ("emp_dt_min", ctypes.c_double), ("emp_dt_max", ctypes.c_double),
    ("bcup_dis", ctypes.c_int32), ("bcup_part", ctypes.c_int32),
    ("bcdown", ctypes.c_int32), ("clamp", ctypes.c_int32),
    ("arith", ctypes.c_int32), ("emp_weights", ctypes.c_int32), ("adaptive_method", ctypes.c_int32),
]


def fill_struct(struct_type, values: dict):
    s = struct_type()
    for name, _ in struct_type._fields_:
        setattr(s, name, values[name])
    return s


def params_dict(p: SedParams | dict) -> dict:
    return p.si() if isinstance(p, SedParams) else dict(p)


__all__ = [
    "VARIABLES", "NVAR", "PARTICULATE", "ELEMENT_VARS", "SedParams", "OmexdiaNamelist",
    "PARAM_FIELDS", "fill_struct", "params_dict", "asdict",
    "EULER", "RUNGE_KUTTA_4", "ADAPTIVE_EULER", "RUNGE_KUTTA_4_38", "EMP1", "EMP2", "EMP_ADAPTIVE", "ROS2_ADAPTIVE",
]
