"""The compile-time e-e-n polynomial of the HIP kernels (champ_amd/csrc/een_poly.hpp),
built for the host against the emulation headers, agrees with the direct term sum of
jastrow4.f90:203-262 in value and first and second derivatives (tests/cpu_emu/een_check.cpp);
and the kernels' branch-free Cartesian monomials agree with slm.f90's case-by-case form."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_een_polynomial_tables(tmp_path):
    exe = tmp_path / "een_check"
    emu = os.path.join(HERE, "cpu_emu")
    subprocess.run(["g++", "-std=c++20", "-O1", "-pthread", "-I" + os.path.join(emu, "include"), "-x", "c++",
                    os.path.join(emu, "een_check.cpp"), "-x", "none", os.path.join(emu, "emu.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_branch_free_monomials(tmp_path):
    """slm_decode / slm_all (the kernels' Cartesian monomials) against the
    case-by-case slm.f90 form for every l (tests/cpu_emu/slm_check.cpp)."""
    exe = tmp_path / "slm_check"
    emu = os.path.join(HERE, "cpu_emu")
    subprocess.run(["g++", "-std=c++20", "-O1", "-pthread", "-I" + os.path.join(emu, "include"), "-x", "c++",
                    os.path.join(emu, "slm_check.cpp"), "-x", "none", os.path.join(emu, "emu.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
