"""Build the gfx950 runtime library libbiocfd_hip.so in-tree with hipcc.

    python bio-cfd_amd/build.py            # compile (always) into bio-cfd_amd/lib/

Also builds the Fortran bindings (fortran/biocfd_hip.f90) and the Fortran host
driver (fortran/biocfd_run.f90 -> lib/biocfd_run) with amdflang.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libbiocfd_hip.so")
OBJ_DIR = os.path.join(HERE, "lib", "obj")

SOURCES = ["runtime.hip", "k_pcor.hip", "k_momentum.hip", "k_search.hip", "k_forcing.hip",
           "k_interp.hip", "k_mixed.hip", "k_move.hip", "k_output.hip", "k_adjoint.hip",
           "k_health.hip"]

# FP64 parity needs no contraction; packed-FP32 instruction forms are kept off.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c",
           os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_SOURCES = ["biocfd_hip.f90", "biocfd_run.f90"]
FORTRAN_EXE = os.path.join(LIB_DIR, "biocfd_run")


def flang() -> str:
    for cand in (os.environ.get("FC_AMD"), "/opt/rocm/bin/amdflang", "/opt/rocm/llvm/bin/flang"):
        if cand and os.path.exists(cand):
            return cand
    return "amdflang"


def build_fortran(verbose: bool = True) -> str:
    """The Fortran bindings module and the Fortran host driver, linked
    against the library (run path $ORIGIN)."""
    mod_dir = os.path.join(OBJ_DIR, "fortran")
    os.makedirs(mod_dir, exist_ok=True)
    objs = []
    for src in FORTRAN_SOURCES:
        obj = os.path.join(mod_dir, os.path.splitext(src)[0] + ".o")
        cmd = [flang(), "-O2", "-c", os.path.join(FORTRAN_DIR, src), "-o", obj, "-module-dir", mod_dir,
               "-I", mod_dir]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {src}:\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    tmp = FORTRAN_EXE + ".tmp"
    cmd = [flang(), *objs, "-o", tmp, "-L", LIB_DIR, "-lbiocfd_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, FORTRAN_EXE)
    if verbose:
        print(f"built {FORTRAN_EXE}")
    return FORTRAN_EXE


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
