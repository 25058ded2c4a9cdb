"""The C-ABI library loads without a GPU and exports every entry point that
include/champ_gpu.h declares (no compute call is made here)."""
import ctypes as C
import os
import re

import pytest

from champ_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            with open(os.path.join(ROOT, "include", fn)) as f:
                names |= set(re.findall(r"\b(champ_gpu_\w+)\s*\(", f.read()))
    return names


def test_header_declares_the_python_binding_set():
    assert set(E.EXPORTS) <= _declared()


def test_library_exports_every_declared_symbol():
    if not os.path.exists(E.LIBPATH):
        pytest.fail("libchamp_gpu.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(E.LIBPATH)
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_init_rejects_bad_arguments_without_touching_the_device():
    lib = E.load_library()
    ptr = C.c_void_p()
    # nwalk <= 0 is refused before any HIP call
    assert lib.champ_gpu_init(C.byref(ptr), None, 0, 0) == -1
    assert b"bad arguments" in lib.champ_gpu_last_error()
