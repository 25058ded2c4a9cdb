"""ctypes access to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of bench.py as the checker / baseline; the engine never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_RNG = os.path.join(HERE, "_ref", "libxoshiro_ref.so")

_lib = None

dp = C.POINTER(C.c_double)
u64p = C.POINTER(C.c_uint64)


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.orc_walker_new.restype = vp
        L.orc_walker_new.argtypes = [vp]
        L.orc_walker_free.argtypes = [vp]
        for f in ("orc_walker_set_x", "orc_walker_get_x"):
            getattr(L, f).argtypes = [vp, dp]
        L.orc_walker_set_quad.argtypes = [vp, dp, dp, dp]
        L.orc_walker_set_rng.argtypes = [vp, u64p]
        L.orc_walker_get_rng.argtypes = [vp, u64p]
        L.orc_hpsi.argtypes = [vp, vp, dp, dp, dp, dp]
        L.orc_psie.argtypes = [vp, vp, C.c_int, dp, dp, dp, C.c_int]
        L.orc_determinante_grad.argtypes = [vp, vp, C.c_int, C.c_int, dp]
        L.orc_accept.argtypes = [vp, vp, C.c_int]
        L.orc_reject.argtypes = [vp, vp, C.c_int]
        L.orc_metrop1_sweep.argtypes = [vp, vp, C.c_double]
        L.orc_metrop1_sweep.restype = C.c_int
        L.orc_metrop6_sweep.argtypes = [vp, vp, C.c_double, C.c_double, C.c_double]
        L.orc_dmc_new.argtypes = [vp, C.c_int, dp, u64p, C.c_int, C.c_double, C.c_double, C.c_int, u64p]
        L.orc_dmc_new.restype = vp
        L.orc_dmc_free.argtypes = [vp]
        L.orc_dmc_set_icasula.argtypes = [vp, C.c_int]
        L.orc_dmc_step.argtypes = [vp, vp, dp]
        L.orc_dmc_step.restype = C.c_int
        L.orc_dmc_block.argtypes = [vp]
        L.orc_dmc_block.restype = C.c_double
        L.orc_dmc_get.argtypes = [vp, dp, dp, dp]
        L.orc_dmc_get.restype = C.c_int
        L.orc_metrop6_sweep.restype = C.c_int
        L.orc_gammai.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int)]
        L.orc_gammai.restype = C.c_double
        L.orc_m6_area.argtypes = [C.c_double] * 5
        L.orc_m6_area.restype = C.c_double
        L.orc_walker_scalars.argtypes = [vp, dp]
        for f in ("orc_get_vj", "orc_get_vjn", "orc_get_slmi", "orc_get_detiab", "orc_get_orbn", "orc_get_parts"):
            getattr(L, f).argtypes = [vp, dp]
        L.orc_get_ddx.argtypes = [vp, dp, dp]
        L.orc_jastrow_full.argtypes = [vp, vp, dp, dp, dp]
        L.orc_powi.restype = C.c_double
        L.orc_powi.argtypes = [C.c_double, C.c_int]
        L.orc_splfit.argtypes = [vp, C.c_int, C.c_int, C.c_double, dp, C.c_int]
        L.orc_slm.argtypes = [C.c_int, dp, C.c_double, dp, dp, dp, C.c_int]
        L.orc_basis_fns_point.argtypes = [vp, dp, dp, C.c_int, dp, dp, dp]
        L.orc_matinv.argtypes = [dp, C.c_int, dp]
        L.orc_matinv.restype = C.c_int
        L.orc_nterms4.argtypes = [C.c_int]
        L.orc_nterms4.restype = C.c_int
        L.orc_asymp.argtypes = [vp, dp, dp]
        L.orc_pot_nn.argtypes = [vp]
        L.orc_pot_nn.restype = C.c_double
        L.orc_gauss_pot.argtypes = [vp, C.c_double, C.c_int, C.c_int, dp, dp]
        L.orc_yl0.argtypes = [C.c_int, C.c_double]
        L.orc_yl0.restype = C.c_double
        L.orc_gesqua.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_xoshiro_next.argtypes = [u64p]
        L.orc_xoshiro_next.restype = C.c_uint64
        L.orc_xoshiro_jump.argtypes = [u64p]
        L.orc_random_dp.argtypes = [u64p]
        L.orc_random_dp.restype = C.c_double
        L.orc_gauss.argtypes = [u64p]
        L.orc_gauss.restype = C.c_double
        L.orc_setrn.argtypes = [u64p, C.POINTER(C.c_int32)]
        L.orc_rotqua.argtypes = [vp, u64p, dp, dp, dp]
        L.orc_vmc_run.argtypes = [vp, C.c_int, dp, u64p, C.c_int, C.c_double, C.c_int, dp, dp]
        L.orc_vmc_run.restype = C.c_double
        _lib = L
    return _lib


def host_cpu():
    """(model name, logical CPUs, CPUs this process may run on, x86-64 level)."""
    model, flags = "unknown", set()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}
    v3 = {"avx", "avx2", "fma", "bmi1", "bmi2", "f16c", "movbe"}
    level = "x86-64-v4" if v4 <= flags and v3 <= flags else "x86-64-v3" if v3 <= flags else "x86-64-v2"
    return model, os.cpu_count() or 1, len(os.sched_getaffinity(0)), level


def baseline_lib():
    """The oracle built for the host's highest x86-64 level (the CPU baseline
    leg of the bench), with orc_vmc_run bound; returns (library, level)."""
    level = host_cpu()[3]
    name = {"x86-64-v4": "liboracle_v4.so", "x86-64-v3": "liboracle_v3.so"}.get(level, "liboracle.so")
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        path, level = LIB, "x86-64-v2"
    L = C.CDLL(path)
    L.orc_vmc_run.argtypes = [C.c_void_p, C.c_int, dp, u64p, C.c_int, C.c_double, C.c_int, dp, dp]
    L.orc_vmc_run.restype = C.c_double
    return L, level


def P(a: np.ndarray):
    return a.ctypes.data_as(dp)


def U(a: np.ndarray):
    return a.ctypes.data_as(u64p)


class Walker:
    """One oracle walker bound to a system descriptor."""

    def __init__(self, csys, nelec: int, norb: int, nup: int):
        self.L = lib()
        self.csys = csys
        self.ptr = self.L.orc_walker_new(C.byref(csys))
        self.nelec, self.norb, self.nup = nelec, norb, nup
        self.nmax = max(nup, nelec - nup)

    def __del__(self):
        try:
            self.L.orc_walker_free(self.ptr)
        except Exception:
            pass

    def set_x(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        self.L.orc_walker_set_x(self.ptr, P(x))

    def get_x(self):
        x = np.zeros((self.nelec, 3))
        self.L.orc_walker_get_x(self.ptr, P(x))
        return x

    def set_quad(self, xq, yq, zq):
        xq, yq, zq = (np.ascontiguousarray(a, dtype=np.float64) for a in (xq, yq, zq))
        self.L.orc_walker_set_quad(self.ptr, P(xq), P(yq), P(zq))

    def set_rng(self, st):
        st = np.ascontiguousarray(st, dtype=np.uint64)
        self.L.orc_walker_set_rng(self.ptr, U(st))

    def get_rng(self):
        st = np.zeros(4, dtype=np.uint64)
        self.L.orc_walker_get_rng(self.ptr, U(st))
        return st

    def hpsi(self):
        o = np.zeros(4)
        self.L.orc_hpsi(C.byref(self.csys), self.ptr, P(o[0:]), P(o[1:]), P(o[2:]), P(o[3:]))
        return dict(psid=o[0], psij=o[1], ekin=o[2], energy=o[3])

    def psie(self, iel, xnew, iflag=0):
        xnew = np.ascontiguousarray(xnew, dtype=np.float64)
        o = np.zeros(2)
        self.L.orc_psie(C.byref(self.csys), self.ptr, int(iel), P(xnew), P(o[0:]), P(o[1:]), int(iflag))
        return o[0], o[1]

    def grad(self, iel, iflag_move):
        v = np.zeros(3)
        self.L.orc_determinante_grad(C.byref(self.csys), self.ptr, int(iel), int(iflag_move), P(v))
        return v

    def accept(self, iel):
        self.L.orc_accept(C.byref(self.csys), self.ptr, int(iel))

    def reject(self, iel):
        self.L.orc_reject(C.byref(self.csys), self.ptr, int(iel))

    def sweep(self, tau):
        return self.L.orc_metrop1_sweep(C.byref(self.csys), self.ptr, float(tau))

    def sweep6(self, deltar=5.0, deltat=1.0, fbias=1.0):
        """one metrop6 sweep (CHAMP's defaults deltar 5, deltat 1, fbias 1)"""
        return self.L.orc_metrop6_sweep(C.byref(self.csys), self.ptr, float(deltar), float(deltat), float(fbias))

    def scalars(self):
        o = np.zeros(4)
        self.L.orc_walker_scalars(self.ptr, P(o))
        return dict(psido=o[0], psijo=o[1], eold=o[2], psi2o=o[3])

    def vj(self):
        o = np.zeros((self.nelec, 3)); self.L.orc_get_vj(self.ptr, P(o)); return o

    def vjn(self):
        o = np.zeros((self.nelec, 3)); self.L.orc_get_vjn(self.ptr, P(o)); return o

    def ddx(self):
        a = np.zeros((self.nelec, 3)); b = np.zeros(self.nelec)
        self.L.orc_get_ddx(self.ptr, P(a), P(b)); return a, b

    def slmi(self):
        o = np.zeros((2, self.nmax, self.nmax)); self.L.orc_get_slmi(self.ptr, P(o)); return o

    def detiab(self):
        o = np.zeros(2); self.L.orc_get_detiab(self.ptr, P(o)); return o

    def orbn(self):
        o = np.zeros((4, self.norb)); self.L.orc_get_orbn(self.ptr, P(o)); return o

    def parts(self):
        o = np.zeros(6); self.L.orc_get_parts(self.ptr, P(o))
        return dict(pe_local=o[0], d2j=o[1], vpsp_up=o[2], vpsp_dn=o[3], nxquad=int(o[4]), d2n=o[5])

    def jastrow_full(self):
        fs = np.zeros(1); d2 = np.zeros(1); fj = np.zeros((self.nelec, 3))
        self.L.orc_jastrow_full(C.byref(self.csys), self.ptr, P(fs), P(d2), P(fj))
        return fs[0], d2[0], fj


def walker(system, csys=None):
    csys = csys if csys is not None else system.as_ctypes()
    w = Walker(csys, system.nelec, system.norb, system.nup)
    w._sys_keep = csys
    return w


def seeded_streams(seed8, nwalk: int) -> np.ndarray:
    """Per-walker xoshiro states: setrn(seed) then jump^w, parser.f90:651-661."""
    L = lib()
    s = np.zeros(4, dtype=np.uint64)
    seed8 = np.ascontiguousarray(seed8, dtype=np.int32)
    L.orc_setrn(U(s), seed8.ctypes.data_as(C.POINTER(C.c_int32)))
    out = np.zeros((nwalk, 4), dtype=np.uint64)
    for w in range(nwalk):
        out[w] = s
        L.orc_xoshiro_jump(U(s))
    return out


def ref_rng_lib():
    """The reference's own xoroshiro256starstar.c built from the CHAMP tree
    (oracle/Makefile target `ref`), or None when absent."""
    if not os.path.exists(REF_RNG):
        return None
    L = C.CDLL(REF_RNG)
    L.next.restype = C.c_uint64
    L.next_uniform_double.restype = C.c_double
    L.xoroshiro_set_state.argtypes = [C.POINTER(C.c_uint32)]
    L.xoroshiro_get_state.argtypes = [C.POINTER(C.c_uint32)]
    return L


class Dmc:
    """The oracle's DMC population (orc_dmc_*): cap walker slots with their own
    streams, the first nconf live, and the population stream of splitj."""

    def __init__(self, system, x0, rng0, nconf, tau, etrial, pop, nfprod=100, csys=None, icasula=0):
        self.L = lib()
        self.sys = system
        self.csys = csys if csys is not None else system.as_ctypes()
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        rng0 = np.ascontiguousarray(rng0, dtype=np.uint64)
        pop = np.ascontiguousarray(pop, dtype=np.uint64)
        self.cap = x0.shape[0]
        self.ptr = self.L.orc_dmc_new(C.byref(self.csys), self.cap, P(x0), U(rng0), int(nconf), float(tau),
                                      float(etrial), int(nfprod), U(pop))
        self.L.orc_dmc_set_icasula(self.ptr, int(icasula))

    def __del__(self):
        if getattr(self, "ptr", None):
            self.L.orc_dmc_free(self.ptr)
            self.ptr = None

    def step(self):
        out = np.zeros(8)
        n = self.L.orc_dmc_step(C.byref(self.csys), self.ptr, P(out))
        if n < 0:
            raise RuntimeError("oracle DMC: walker capacity exceeded")
        keys = ("wsum", "esum", "wgsum", "egsum", "eest", "eigv", "fprod", "nodecr")
        d = dict(zip(keys, out))
        d["nwalk"] = n
        return d

    def block(self):
        return self.L.orc_dmc_block(self.ptr)

    def population(self):
        ne = self.sys.nelec
        wt = np.zeros(self.cap)
        x = np.zeros((self.cap, ne, 3))
        e = np.zeros(self.cap)
        n = self.L.orc_dmc_get(self.ptr, P(wt), P(x), P(e))
        return wt[:n], x[:n], e[:n]
