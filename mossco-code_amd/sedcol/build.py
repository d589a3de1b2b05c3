"""Build recipe for libsedcol.so (gfx950) and the Fortran binding.

Run through ``__graft_entry__.build()`` or ``python -m sedcol.build`` (with
mossco-code_amd/ on sys.path).  Outputs stay in-tree so that they travel
with the repository to the GPU machine.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent          # mossco-code_amd/sedcol
ROOT = PKG.parent                              # mossco-code_amd
REPO = ROOT.parent
CSRC = ROOT / "csrc"
FSRC = ROOT / "fortran"
INCLUDE = REPO / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("SEDCOL_FC", "/opt/rocm/bin/amdflang")
ARCH = os.environ.get("SEDCOL_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-shared",
    # FP64 results must follow the written operation order (fused
    # multiply-adds are explicit fma() calls)
    "-ffp-contract=off",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def build_library(out: Path | None = None, extra=()) -> Path:
    out = Path(out) if out else PKG / "libsedcol.so"
    srcs = [CSRC / "sedcol_kernels.hip", CSRC / "sedcol_api.hip"]
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, *extra, f"-I{INCLUDE}", f"-I{CSRC}", *srcs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_surrogate(out: Path | None = None, extra=()) -> Path:
    """libsedsurr.so: the spin-up surrogate's training and inference kernels."""
    out = Path(out) if out else PKG / "libsedsurr.so"
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, *extra, f"-I{INCLUDE}", f"-I{CSRC}", CSRC / "sedsurr.hip", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_fortran(lib: Path | None = None) -> Path | None:
    """ISO_C_BINDING module + the fabmsed-style driver program."""
    if not shutil.which(FC) and not Path(FC).exists():
        print(f"{FC} not found: Fortran binding not built", flush=True)
        return None
    lib = Path(lib) if lib else PKG / "libsedcol.so"
    bdir = FSRC / "_build"
    bdir.mkdir(exist_ok=True)
    _run([FC, "-O2", "-ffp-contract=off", "-c", FSRC / "sedcol_hip.F90", "-o", bdir / "sedcol_hip.o"], cwd=bdir)
    exe = FSRC / "fabmsed_hip"
    rel = os.path.relpath(lib.parent, FSRC)
    _run([FC, "-O2", "-ffp-contract=off", "-I", bdir, FSRC / "fabmsed_hip.F90", bdir / "sedcol_hip.o",
          "-o", exe, f"-L{lib.parent}", "-lsedcol", f"-Wl,-rpath,$ORIGIN/{rel}"], cwd=bdir)
    return exe


def main(argv=None):
    build_library()
    build_surrogate()
    if (FSRC / "sedcol_hip.F90").exists():
        build_fortran()


if __name__ == "__main__":
    sys.exit(main())
