"""Build the engine's shared library (libchamp_gpu.so) for gfx950.

hipcc cross-compiles the device code without a GPU; the library depends on
the HIP runtime only (every contraction is an engine kernel).  Run as
``python -m champ_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libchamp_gpu.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["engine_api.hip", "engine_kernels.hip", "engine_hpsi.hip", "engine_gemm.hip", "engine_nonloc.hip",
           "engine_multidet.hip", "engine_dmc.hip", "engine_deriv.hip"]
ARCH = os.environ.get("CHAMP_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # never emit packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: str, objdir: str = OBJDIR, defines: tuple = ()) -> str:
    obj = os.path.join(objdir, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = True, out: str = OUT, defines: tuple = ()) -> str:
    """Compile and link the library.  ``defines`` set the kernels' tuning
    macros (CHAMP_MJB, CHAMP_MOVE_WPE, CHAMP_HJB, CHAMP_HJAS_WPE, CHAMP_NL_QB,
    CHAMP_NL_WPE); ``out`` other than the default builds a variant elsewhere."""
    objdir = OBJDIR if not defines and out == OUT else out + ".objs"
    os.makedirs(objdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda src: _compile(src, objdir, tuple(defines)), SOURCES))
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


FDIR = os.path.join(HERE, "fortran")
FEXE = os.path.join(FDIR, "vmc_gpu")


def build_fortran(verbose: bool = True, out: str = FEXE) -> str:
    """The Fortran binding module and population VMC driver (amdflang), linked
    against libchamp_gpu.so in this package (found at run time through an
    rpath relative to the executable when it is built in place)."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    moddir = os.path.join(OBJDIR, "fmod") if out == FEXE else out + ".mod"
    os.makedirs(moddir, exist_ok=True)
    rpath = "$ORIGIN/.." if out == FEXE else HERE
    cmd = [fc, "-O2", "-J", moddir, os.path.join(FDIR, "champ_gpu_mod.f90"), os.path.join(FDIR, "vmc_gpu.f90"),
           "-L", HERE, "-lchamp_gpu", f"-Wl,-rpath,{rpath}", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stderr[-4000:]}")
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    try:
        build()
        build_fortran()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
