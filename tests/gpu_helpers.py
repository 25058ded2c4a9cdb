"""Shared helpers for the GPU parity tests: one population on the engine and
the same walkers on the CPU oracle."""
import os

import numpy as np

from champ_amd import system as S
from champ_amd.engine import LIBPATH, Engine
from oracle import ffi

# CHAMP_ENGINE_LIB selects another build of the same C-ABI (the CPU emulation
# build under tests/cpu_emu runs these tests on the host).
ENGINE_LIB = os.environ.get("CHAMP_ENGINE_LIB", LIBPATH)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den))


def relnorm(a, b):
    """max |a-b| / max |b| over the array (for vectors with near-zero entries)"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


class Pair:
    """Engine population + matching oracle walkers."""

    def __init__(self, name: str, nwalk: int, seed: int = 11):
        self.sys = S.config(name)
        self.csys = self.sys.as_ctypes()
        self.nwalk = nwalk
        self.x = self.sys.initial_walkers(nwalk, seed=seed)
        self.eng = Engine(self.sys, nwalk, lib_path=ENGINE_LIB)
        self.eng.set_walkers(self.x)
        self.orc = [ffi.walker(self.sys, self.csys) for _ in range(nwalk)]
        for w in range(nwalk):
            self.orc[w].set_x(self.x[w])

    def quad(self, seed: int = 3):
        """Same rotated quadrature on both sides (one rotation per walker)."""
        if self.sys.nloc == 0:
            return
        L = ffi.lib()
        streams = ffi.seeded_streams(S.seed_from_string(), self.nwalk)
        nq = len(self.sys.wq)
        X = np.zeros((self.nwalk, nq)); Y = np.zeros_like(X); Z = np.zeros_like(X)
        for w in range(self.nwalk):
            st = streams[w].copy()
            L.orc_rotqua(__import__("ctypes").byref(self.csys), ffi.U(st), ffi.P(X[w]), ffi.P(Y[w]), ffi.P(Z[w]))
            self.orc[w].set_quad(X[w], Y[w], Z[w])
        self.eng.set_quadrature(X, Y, Z)
