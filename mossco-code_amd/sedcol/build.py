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
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC",
    # FP64 results must follow the written operation order (fused
    # multiply-adds are explicit fma() calls)
    "-ffp-contract=off",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


LIB_SOURCES = [CSRC / "sedcol_kernels.hip", CSRC / "sedcol_api.hip"]
SURR_SOURCES = [CSRC / "sedsurr.hip"]
OBJ_DIR = PKG / "_build"


def _compile(srcs, extra=()) -> list[Path]:
    """Compile HIP sources to objects, all at once in separate processes."""
    OBJ_DIR.mkdir(exist_ok=True)
    jobs = []
    for src in srcs:
        obj = OBJ_DIR / (Path(src).stem + ".o")
        cmd = [HIPCC, *HIP_FLAGS, *extra, f"-I{INCLUDE}", f"-I{CSRC}", "-c", src, "-o", obj]
        print("+", " ".join(str(c) for c in cmd), flush=True)
        jobs.append((subprocess.Popen([str(c) for c in cmd]), cmd, obj))
    failed = [cmd for proc, cmd, _ in jobs if proc.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, [str(c) for c in failed[0]])
    return [obj for _, _, obj in jobs]


def _link(objs, out: Path) -> Path:
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_library(out: Path | None = None, extra=()) -> Path:
    return _link(_compile(LIB_SOURCES, extra), Path(out) if out else PKG / "libsedcol.so")


def build_surrogate(out: Path | None = None, extra=()) -> Path:
    """libsedsurr.so: the spin-up surrogate's training and inference kernels."""
    return _link(_compile(SURR_SOURCES, extra), Path(out) if out else PKG / "libsedsurr.so")


def build_all(extra=()) -> tuple[Path, Path]:
    """Both libraries, with every source compiled in parallel."""
    objs = _compile(LIB_SOURCES + SURR_SOURCES, extra)
    lib = _link(objs[: len(LIB_SOURCES)], PKG / "libsedcol.so")
    surr = _link(objs[len(LIB_SOURCES):], PKG / "libsedsurr.so")
    return lib, surr


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
    for prog in ("fabmsed_hip", "fabmsed1d_hip"):
        _run([FC, "-O2", "-ffp-contract=off", "-I", bdir, FSRC / f"{prog}.F90", bdir / "sedcol_hip.o",
              "-o", FSRC / prog, f"-L{lib.parent}", "-lsedcol", f"-Wl,-rpath,$ORIGIN/{rel}"], cwd=bdir)
    return exe


def main(argv=None):
    build_all()
    if (FSRC / "sedcol_hip.F90").exists():
        build_fortran()


if __name__ == "__main__":
    sys.exit(main())
