"""The Fortran side of the boundary (champ_amd/fortran): the iso_c_binding
module compiles against the library, its bind(c) types have the layout of
the C declarations in include/champ_gpu.h, every bind(c) name is a symbol the
library exports, and (on the GPU) the Fortran population driver reproduces
the Python driver's sweeps through the same C-ABI.

CHAMP binds its C random-number file the same way (src/vmc/random.f90:3-22)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from champ_amd import build as B
from champ_amd import system as S
from champ_amd.engine import LIBPATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOD = os.path.join(ROOT, "champ_amd", "fortran", "champ_gpu_mod.f90")
HDR = os.path.join(ROOT, "include", "champ_gpu.h")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


def _layout(cmd, exe):
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return dict(line.split() for line in out.strip().splitlines())


@pytest.mark.skipif(not os.path.exists(LIBPATH), reason="engine library not built")
def test_bind_c_types_match_the_c_layout(tmp_path):
    """Offsets of every component of champ_system_t, champ_sweep_stats,
    champ_vmc_result, champ_dmc_step_stats, champ_dmc_result and
    champ_sr_info, and their sizes, as amdflang lays out the bind(c) types and as the C
    compiler lays out the header's structs."""
    fexe, cexe = str(tmp_path / "lf"), str(tmp_path / "lc")
    f = _layout([FC, "-O0", "-J", str(tmp_path), MOD, os.path.join(ROOT, "tests", "fortran", "layout_check.f90"),
                 "-L", os.path.dirname(LIBPATH), "-lchamp_gpu", f"-Wl,-rpath,{os.path.dirname(LIBPATH)}",
                 "-o", fexe], fexe)
    c = _layout(["gcc", "-I", os.path.dirname(HDR), os.path.join(ROOT, "tests", "fortran", "layout_check.c"),
                 "-o", cexe], cexe)
    assert len(c) == 54 + 8 + 13 + 10 + 5
    assert f == c


def _bind_names():
    return re.findall(r'bind\(c,\s*name="([a-z0-9_]+)"\)', open(MOD).read())


@pytest.mark.skipif(not os.path.exists(LIBPATH), reason="engine library not built")
def test_bind_c_names_are_exported_and_declared():
    names = _bind_names()
    assert len(names) >= 26
    syms = subprocess.run(["nm", "-D", "--defined-only", LIBPATH], check=True, capture_output=True,
                          text=True).stdout.split()
    hdr = open(HDR).read()
    for n in names:
        assert n in syms, n
        assert re.search(rf"\b{n}\(", hdr), n


@pytest.mark.skipif(not os.path.exists(LIBPATH), reason="engine library not built")
def test_fortran_driver_builds(tmp_path):
    """The module and the population driver compile and link (amdflang)."""
    exe = B.build_fortran(verbose=False, out=str(tmp_path / "vmc_gpu"))
    assert os.access(exe, os.X_OK)


@pytest.mark.gpu
def test_fortran_driver_matches_python_driver(require_gpu, tmp_path):
    """vmc_gpu (Fortran, through champ_gpu_mod) and Engine (Python, ctypes)
    on the same system file, walkers and seed: the block energies of their
    metrop1 and metrop6 sweeps agree to 1e-13 relative (the walkers follow the same
    chains; the population's energy sum is accumulated with atomics, whose
    order, and so the last bits of the sum, vary from run to run)."""
    if os.environ.get("CHAMP_ENGINE_LIB"):
        pytest.skip("the Fortran driver links the GPU library itself")
    from champ_amd.engine import Engine
    if not os.access(B.FEXE, os.X_OK):
        raise RuntimeError(f"{B.FEXE} is not built: run __graft_entry__.build() first")
    for name, nwalk, imetro in (("small", 16, 1), ("c22", 8, 1), ("small", 16, 6)):
        sy = S.config(name)
        x0 = sy.initial_walkers(nwalk, seed=21)
        path = str(tmp_path / f"{name}.bin")
        sy.write_binary(path, x0)
        nblk, nstep, tau = 3, 2, 0.3
        print(f"[fortran driver] {name} nwalk={nwalk} imetro={imetro}", flush=True)
        r = subprocess.run([B.FEXE, path, str(nblk), str(nstep), str(tau), "18123437465549275475255231234865",
                            str(imetro)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        ef = [float(m) for m in re.findall(r"block\s+\d+\s+E_L\s+(\S+)", r.stdout)]
        assert len(ef) == nblk, r.stdout
        e = Engine(sy, nwalk)
        e.set_walkers(x0)
        e.seed_rng(S.seed_from_string())
        e.hpsi()
        if imetro == 6:
            ep = [e.metrop6(nstep)["esum"] / (nwalk * nstep) for _ in range(nblk)]
        else:
            ep = [e.metrop1(nstep, tau)["esum"] / (nwalk * nstep) for _ in range(nblk)]
        print(f"[fortran driver] {name} fortran {ef} python {ep}", flush=True)
        np.testing.assert_allclose(ef, ep, rtol=1e-13, atol=0)
