"""Algorithmic FP64 operation counts of the hot path.

Counted once from the formulas of mossco-code_amd/csrc/sedcol_device.hpp and
sedcol_kernels.hip (add, sub, mul, div = 1; fma = 2; sign flips, min/max,
comparisons and selects = 0).  Division and the per-column exp count as one
operation each, so issued instructions exceed these figures (an FP64 divide
expands to about ten VALU instructions).

Per cell:
  kinetics()                       66   (detritus 16, pathway split 21,
                                         nitrification/ODU/P 17, rates 12)
  transport per particulate var    12   (x=c*phi 1, F 3, VF 2, divergence 3,
                                         rescale 3, +R 1 ... counted below)
  transport per dissolved var       7   (F 3, divergence 3, +R 1)
  rhs = kinetics + 6*12 + 5*7   = 173
Adaptive EMP attempt (two stages, per cell):
  2 rhs                           346
  Patankar set-up, 2 x 11 x 2      44   (h*g, /c)
  second-stage scale, 11 x 2       22   (c/c1, product)
  stage c1 = c + (h p1) f0         22
  g = (f0 + f1) * 0.5              22
  error norm 11 x 6                66
  accepted update 11 x 2 + sinks   26
Newton iteration (per cell):       70   (11 x (fma 2 + 3 mul + 1 add) + 4 for
                                         one (P, D) tree combine)
RK4 step (per cell): 4 rhs + stage updates 3 x 11 x 2 + combination
  11 x (2 + 2) + sinks 12  = 692 + 66 + 44 + 12 = 814.
Linearly implicit ROS2 attempt (method 7, per cell):
  2 rhs                           346
  reaction Jacobian kin_jac, x 2  254   (127 each: detritus 41, partition 59,
                                         nitrification/ODU/P 27)
  transport factor, 11 species    165   (15 per species and layer)
  2 W-solves                     1578   (789 each: 2 species sweeps 2 x 132,
                                         2 block solves 2 x 174, J_R k0 67,
                                         transport residual 99, sum 11)
  stage input and f1 - 2 k1        44
  sink gradients                   24
  error norm 11 x 13 + 1          144
  accepted update + sinks          65
  total                          2620
"""
from __future__ import annotations

import numpy as np

NV = 11
N_PART, N_DIS = 6, 5
KINETICS = 16 + 21 + 17 + 12
TRANSPORT_PART = 12
TRANSPORT_DIS = 7
RHS_PER_CELL = KINETICS + N_PART * TRANSPORT_PART + N_DIS * TRANSPORT_DIS  # 173
EMP_ATTEMPT_PER_CELL = 2 * RHS_PER_CELL + 2 * NV * 2 + NV * 2 + NV * 2 + NV * 2 + NV * 6 + (NV * 2 + 4)
NEWTON_PER_CELL = NV * 6 + 4
#: geometric-mean weights: the scaled factors and their derivative, a log
#: (13-term series, ~30 operations), the derivative ratio and two sums
NEWTON_GEO_PER_CELL = NV * 7 + 36
RK4_STEP_PER_CELL = 4 * RHS_PER_CELL + 3 * NV * 2 + NV * 4 + 12
KIN_JAC = 41 + 59 + 27
KJAC_SOLVE = 174
KJAC_APPLY = 67
SPECIES_SWEEP = 12 * NV
W_SOLVE = 2 * SPECIES_SWEEP + 2 * KJAC_SOLVE + KJAC_APPLY + 9 * NV + NV
ROS2_ATTEMPT_PER_CELL = (2 * RHS_PER_CELL + 2 * KIN_JAC + 15 * NV + 2 * W_SOLVE + 4 * NV + 24
                         + (13 * NV + 1) + (5 * NV + 10))  # 2620

#: HBM bytes a persistent adaptive launch moves per cell and per column,
#: independent of the number of steps inside it: state in+out (2*11*8), the
#: DIC and N2 sink sums with their compensations in+out (2*4*8), geometry
#: (dz, dzc, phi, phiI, bfac: 5*8); per column: offsets, bioturbation, top
#: flux, its time integral and that integral's compensation in+out
#: (2*3*11*8), diagnostics in+out, and per forcing record temperature +
#: 2*11 forcing values.
BYTES_PER_CELL_LAUNCH = 2 * NV * 8 + 2 * 4 * 8 + 5 * 8
BYTES_PER_COLUMN_LAUNCH = 8 + 8 + 2 * 3 * NV * 8 + 2 * 88
BYTES_PER_COLUMN_RECORD = 8 + 2 * NV * 8


def emp_flops(cells_per_col, attempts, newton, weights=0):
    """FLOPs of adaptive-EMP work: attempts and Newton iterations per column
    (weights: the Patankar weighting, 0 product, 1 geometric mean)."""
    K = np.asarray(cells_per_col, dtype=np.float64)
    per_newton = NEWTON_GEO_PER_CELL if weights == 1 else NEWTON_PER_CELL
    return float(np.sum(K * (EMP_ATTEMPT_PER_CELL * np.asarray(attempts, dtype=np.float64)
                             + per_newton * np.asarray(newton, dtype=np.float64))))


def ros2_flops(cells_per_col, attempts):
    """FLOPs of linearly implicit (method 7) work: attempts per column."""
    K = np.asarray(cells_per_col, dtype=np.float64)
    return float(np.sum(K * ROS2_ATTEMPT_PER_CELL * np.asarray(attempts, dtype=np.float64)))


def rk4_flops(cells_per_col, steps):
    K = np.asarray(cells_per_col, dtype=np.float64)
    return float(np.sum(K * RK4_STEP_PER_CELL * np.asarray(steps, dtype=np.float64)))


def launch_bytes(cells_per_col, records=1):
    K = np.asarray(cells_per_col, dtype=np.float64)
    return float(np.sum(K * BYTES_PER_CELL_LAUNCH + BYTES_PER_COLUMN_LAUNCH + records * BYTES_PER_COLUMN_RECORD))
