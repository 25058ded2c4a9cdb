"""Python host side of the engine: ctypes bindings of include/champ_gpu.h.

The methods keep the names and argument meaning of the CHAMP routines they
stand for (hpsi, psie, compute_determinante_grad, jassav/detsav via accept,
metrop1, rotqua), each acting on a whole walker population.  There is no CPU
fallback: constructing an :class:`Engine` without the built library or
without a GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

from .system import CSystem, System

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libchamp_gpu.so")

PREC_FP64 = 0
PREC_MIXED = 1

ARR = dict(slmi=1, vj=2, ddx=3, d2dx2=4, detiab=5, vpsp=6, peloc=7, d2j=8, orbn=9, vjn=10, nxquad=11, fso=12)

EXPORTS = [
    "champ_gpu_init", "champ_gpu_free", "champ_gpu_last_error", "champ_gpu_set_walkers", "champ_gpu_get_walkers",
    "champ_gpu_set_rng", "champ_gpu_get_rng", "champ_gpu_seed_rng", "champ_gpu_set_quadrature", "champ_gpu_rotqua",
    "champ_gpu_hpsi", "champ_gpu_compute_determinante_grad", "champ_gpu_psie", "champ_gpu_accept",
    "champ_gpu_metrop1", "champ_gpu_get_walker_scalars", "champ_gpu_get_status", "champ_gpu_array_size",
    "champ_gpu_get_array", "champ_gpu_basis_fns", "champ_gpu_kernel_times", "champ_gpu_set_graphs",
    "champ_gpu_autotune", "champ_gpu_metrop1_batch", "champ_gpu_determinant", "champ_gpu_set_lu_every",
    "champ_gpu_set_tuning", "champ_gpu_vmc",
]


class VmcResult(C.Structure):
    _fields_ = [("efin", C.c_double), ("eerr", C.c_double), ("sigma", C.c_double), ("tcorr", C.c_double),
                ("accept", C.c_double), ("nsamples", C.c_int64), ("nflagged", C.c_int64)]


DERIV_JAS_EN, DERIV_JAS_EE, DERIV_JAS_EEN, DERIV_CI, DERIV_MAPPED, DERIV_ORB = 1, 2, 4, 8, 16, 32


class SrInfo(C.Structure):
    _fields_ = [("adiag", C.c_double), ("dparm_norm", C.c_double), ("cg_iters", C.c_int32), ("retries", C.c_int32)]


class DmcStepStats(C.Structure):
    _fields_ = [("nwalk", C.c_int32), ("pad", C.c_int32), ("wsum", C.c_double), ("esum", C.c_double),
                ("wgsum", C.c_double), ("egsum", C.c_double), ("eest", C.c_double), ("eigv", C.c_double),
                ("fprod", C.c_double), ("accept", C.c_double), ("nodecr", C.c_int64), ("nbranch", C.c_int64)]


class DmcResult(C.Structure):
    _fields_ = [("egave", C.c_double), ("egerr", C.c_double), ("eest", C.c_double), ("accept", C.c_double),
                ("tau_eff", C.c_double), ("nwalk", C.c_int32), ("nwalk_max", C.c_int32), ("nbranch", C.c_int64),
                ("nodecr", C.c_int64)]


class SweepStats(C.Structure):
    _fields_ = [("accepted", C.c_double), ("proposed", C.c_double), ("esum", C.c_double), ("esum2", C.c_double),
                ("nflagged", C.c_int64)]


_libs: Dict[str, C.CDLL] = {}


def load_library(path: str = LIBPATH):
    """Load libchamp_gpu.so (or another build of the same C-ABI); raises if it
    has not been built."""
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m champ_amd.build`")
    L = C.CDLL(path)
    vp, dp, i32 = C.c_void_p, C.POINTER(C.c_double), C.c_int32
    L.champ_gpu_init.argtypes = [C.POINTER(vp), C.POINTER(CSystem), i32, i32]
    L.champ_gpu_free.argtypes = [vp]
    L.champ_gpu_last_error.restype = C.c_char_p
    L.champ_gpu_set_walkers.argtypes = [vp, dp]
    L.champ_gpu_get_walkers.argtypes = [vp, dp]
    L.champ_gpu_set_rng.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.champ_gpu_get_rng.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.champ_gpu_seed_rng.argtypes = [vp, C.POINTER(C.c_int32)]
    L.champ_gpu_set_quadrature.argtypes = [vp, dp, dp, dp]
    L.champ_gpu_rotqua.argtypes = [vp]
    L.champ_gpu_hpsi.argtypes = [vp, dp, dp, dp, dp]
    L.champ_gpu_compute_determinante_grad.argtypes = [vp, i32, i32, dp]
    L.champ_gpu_psie.argtypes = [vp, i32, dp, dp, dp, i32]
    L.champ_gpu_accept.argtypes = [vp, i32, C.POINTER(C.c_uint8)]
    L.champ_gpu_metrop1.argtypes = [vp, i32, C.c_double, C.POINTER(SweepStats)]
    L.champ_gpu_metrop6.argtypes = [vp, i32, C.POINTER(SweepStats)]
    L.champ_gpu_dmc_init.argtypes = [vp, i32, C.c_double, C.c_double, i32, C.POINTER(C.c_uint64)]
    L.champ_gpu_dmc_step.argtypes = [vp, C.POINTER(DmcStepStats)]
    L.champ_gpu_dmc.argtypes = [vp, i32, i32, i32, dp, C.POINTER(DmcResult)]
    L.champ_gpu_dmc_weights.argtypes = [vp, dp, C.POINTER(i32)]
    L.champ_gpu_dmc_set_icasula.argtypes = [vp, i32]
    L.champ_gpu_deriv_count.argtypes = [vp, i32, C.POINTER(i32)]
    L.champ_gpu_deriv_psi.argtypes = [vp, i32, dp]
    L.champ_gpu_sr_sums.argtypes = [vp, i32, dp, dp, dp, dp]
    L.champ_gpu_set_deriv_map.argtypes = [vp, i32, i32, dp]
    L.champ_gpu_set_params.argtypes = [vp, dp, dp, dp, dp]
    L.champ_gpu_set_orb_terms.argtypes = [vp, i32, C.POINTER(i32)]
    L.champ_gpu_set_orbitals.argtypes = [vp, dp]
    L.champ_sr_solve.argtypes = [i32, C.c_double, dp, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double, dp, C.POINTER(SrInfo)]
    L.champ_gpu_sr_accumulate.argtypes = [vp, i32, i32, dp, dp]
    L.champ_gpu_sr_solve.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, dp, C.POINTER(SrInfo)]
    L.champ_gpu_set_metrop6.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    L.champ_gpu_get_walker_scalars.argtypes = [vp, dp, dp, dp, dp]
    L.champ_gpu_get_status.argtypes = [vp, C.POINTER(C.c_int32)]
    L.champ_gpu_array_size.argtypes = [vp, i32]
    L.champ_gpu_array_size.restype = C.c_int64
    L.champ_gpu_get_array.argtypes = [vp, i32, dp]
    L.champ_gpu_basis_fns.argtypes = [vp, i32, dp, dp]
    L.champ_gpu_basis_fns_noalloc.argtypes = [vp, i32, dp, dp]
    L.champ_gpu_kernel_times.argtypes = [vp, C.POINTER(i32), C.c_char_p, dp, C.POINTER(C.c_int64)]
    L.champ_gpu_set_graphs.argtypes = [vp, i32]
    L.champ_gpu_autotune.argtypes = [vp, C.POINTER(i32)]
    L.champ_gpu_metrop1_batch.argtypes = [C.POINTER(vp), i32, i32, C.c_double, C.POINTER(SweepStats)]
    L.champ_gpu_determinant.argtypes = [vp, dp]
    L.champ_gpu_set_lu_every.argtypes = [vp, i32]
    L.champ_gpu_set_tuning.argtypes = [vp, C.POINTER(i32)]
    L.champ_gpu_vmc.argtypes = [vp, i32, i32, i32, C.c_double, i32, dp, C.POINTER(VmcResult)]
    _libs[path] = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def sr_solve(s, osum, eosum, esum: float, wsum: float, sr_tau: float = 0.02, sr_eps: float = 1e-3,
             sr_adiag: float = 0.01, dparm_norm_min: float = 1.0, lib_path: str = LIBPATH):
    """The SR step from the population sums (champ_sr_solve, host only):
    returns (dparm, info); new parameters are p - dparm."""
    L = load_library(lib_path)
    s = np.ascontiguousarray(s, dtype=np.float64)
    osum = np.ascontiguousarray(osum, dtype=np.float64)
    eosum = np.ascontiguousarray(eosum, dtype=np.float64)
    n = len(osum)
    if s.shape != (n, n) or eosum.shape != (n,):
        raise ValueError("sr_solve: s must be [n][n] and osum, eosum [n]")
    dparm = np.zeros(n)
    info = SrInfo()
    rc = L.champ_sr_solve(n, float(wsum), _p(s), _p(osum), _p(eosum), float(esum), float(sr_tau), float(sr_eps),
                          float(sr_adiag), float(dparm_norm_min), _p(dparm), C.byref(info))
    if rc != 0:
        raise RuntimeError(f"sr_solve failed ({rc}): {L.champ_gpu_last_error().decode()}")
    return dparm, dict(adiag=info.adiag, dparm_norm=info.dparm_norm, cg_iters=info.cg_iters, retries=info.retries)


class Engine:
    """A walker population of one system resident on the current GPU."""

    def __init__(self, system: System, nwalk: int, precision: int = PREC_FP64, lib_path: str = LIBPATH):
        self.L = load_library(lib_path)
        self.system = system
        self.nwalk = int(nwalk)
        self.csys = system.as_ctypes()
        self.ptr = C.c_void_p()
        rc = self.L.champ_gpu_init(C.byref(self.ptr), C.byref(self.csys), self.nwalk, int(precision))
        self._check(rc, "champ_gpu_init")
        self.nelec, self.norb = system.nelec, system.norb
        self.nmax = max(system.nup, system.ndn)

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.L.champ_gpu_last_error().decode()}")

    def close(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            self.L.champ_gpu_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- state
    def set_walkers(self, x: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.nwalk, self.nelec, 3)
        self._check(self.L.champ_gpu_set_walkers(self.ptr, _p(x)), "set_walkers")

    def get_walkers(self) -> np.ndarray:
        x = np.zeros((self.nwalk, self.nelec, 3))
        self._check(self.L.champ_gpu_get_walkers(self.ptr, _p(x)), "get_walkers")
        return x

    def seed_rng(self, seed8):
        s = np.ascontiguousarray(seed8, dtype=np.int32)
        self._check(self.L.champ_gpu_seed_rng(self.ptr, s.ctypes.data_as(C.POINTER(C.c_int32))), "seed_rng")

    def set_rng(self, st: np.ndarray):
        st = np.ascontiguousarray(st, dtype=np.uint64).reshape(self.nwalk, 4)
        self._check(self.L.champ_gpu_set_rng(self.ptr, st.ctypes.data_as(C.POINTER(C.c_uint64))), "set_rng")

    def get_rng(self) -> np.ndarray:
        st = np.zeros((self.nwalk, 4), dtype=np.uint64)
        self._check(self.L.champ_gpu_get_rng(self.ptr, st.ctypes.data_as(C.POINTER(C.c_uint64))), "get_rng")
        return st

    def set_quadrature(self, xq, yq, zq):
        a = [np.ascontiguousarray(np.broadcast_to(v, (self.nwalk, len(self.system.wq))), dtype=np.float64)
             for v in (xq, yq, zq)]
        self._check(self.L.champ_gpu_set_quadrature(self.ptr, _p(a[0]), _p(a[1]), _p(a[2])), "set_quadrature")

    def rotqua(self):
        self._check(self.L.champ_gpu_rotqua(self.ptr), "rotqua")

    # ------------------------------------------------------- CHAMP routines
    def hpsi(self) -> Dict[str, np.ndarray]:
        o = [np.zeros(self.nwalk) for _ in range(4)]
        self._check(self.L.champ_gpu_hpsi(self.ptr, *map(_p, o)), "hpsi")
        return dict(psid=o[0], psij=o[1], ekin=o[2], energy=o[3])

    def determinant(self) -> np.ndarray:
        """FP64 recompute of orbitals, Slater inverse and determinants at the
        current positions (determinant.f90:9-140); returns psid per walker."""
        pd = np.zeros(self.nwalk)
        self._check(self.L.champ_gpu_determinant(self.ptr, _p(pd)), "determinant")
        return pd

    def set_lu_every(self, nmoves: int):
        """Moves between the FP64 recomputes of the inverse inside a metrop1 sweep (0: only in hpsi)."""
        self._check(self.L.champ_gpu_set_lu_every(self.ptr, int(nmoves)), "set_lu_every")

    def compute_determinante_grad(self, iel: int, iflag_move: int) -> np.ndarray:
        v = np.zeros((self.nwalk, 3))
        self._check(self.L.champ_gpu_compute_determinante_grad(self.ptr, int(iel), int(iflag_move), _p(v)),
                    "compute_determinante_grad")
        return v

    def psie(self, iel: int, xnew: np.ndarray, iflag: int = 0):
        xnew = np.ascontiguousarray(xnew, dtype=np.float64).reshape(self.nwalk, 3)
        pd, pj = np.zeros(self.nwalk), np.zeros(self.nwalk)
        self._check(self.L.champ_gpu_psie(self.ptr, int(iel), _p(xnew), _p(pd), _p(pj), int(iflag)), "psie")
        return pd, pj

    def accept(self, iel: int, mask: np.ndarray):
        m = np.ascontiguousarray(mask, dtype=np.uint8).reshape(self.nwalk)
        self._check(self.L.champ_gpu_accept(self.ptr, int(iel), m.ctypes.data_as(C.POINTER(C.c_uint8))), "accept")

    def metrop1(self, nsweep: int, tau: float = 0.5) -> Dict[str, float]:
        st = SweepStats()
        self._check(self.L.champ_gpu_metrop1(self.ptr, int(nsweep), float(tau), C.byref(st)), "metrop1")
        return dict(accepted=st.accepted, proposed=st.proposed, esum=st.esum, esum2=st.esum2, nflagged=st.nflagged)

    def metrop6(self, nsweep: int) -> Dict[str, float]:
        """nsweep metrop6 sweeps (CHAMP's accelerated Metropolis mover)"""
        st = SweepStats()
        self._check(self.L.champ_gpu_metrop6(self.ptr, int(nsweep), C.byref(st)), "metrop6")
        return dict(accepted=st.accepted, proposed=st.proposed, esum=st.esum, esum2=st.esum2, nflagged=st.nflagged)

    # ---------------------------------------------------------------- DMC
    def dmc_init(self, nconf: int, tau: float, etrial: float, pop_state, nfprod: int = 100):
        """DMC population of the first nconf walkers (the context's walker count is the capacity)"""
        st = np.ascontiguousarray(pop_state, dtype=np.uint64).reshape(4)
        self._check(self.L.champ_gpu_dmc_init(self.ptr, int(nconf), float(tau), float(etrial), int(nfprod),
                                              st.ctypes.data_as(C.POINTER(C.c_uint64))), "dmc_init")

    def dmc_set_icasula(self, icasula: int):
        """0: locality approximation; -1: one T-move per walker after each step"""
        self._check(self.L.champ_gpu_dmc_set_icasula(self.ptr, int(icasula)), "dmc_set_icasula")

    def dmc_step(self) -> Dict[str, float]:
        s = DmcStepStats()
        self._check(self.L.champ_gpu_dmc_step(self.ptr, C.byref(s)), "dmc_step")
        return {f: getattr(s, f) for f, _ in DmcStepStats._fields_ if f != "pad"}

    def dmc(self, nblk: int, nstep: int, nblkeq: int = 0) -> Dict[str, float]:
        r = DmcResult()
        eg = np.zeros(int(nblk))
        self._check(self.L.champ_gpu_dmc(self.ptr, int(nblk), int(nstep), int(nblkeq), _p(eg), C.byref(r)), "dmc")
        out = {f: getattr(r, f) for f, _ in DmcResult._fields_}
        out["egnow"] = eg
        return out

    def dmc_population(self):
        """(weights, positions, local energies) of the live walkers"""
        n = C.c_int32()
        wt = np.zeros(self.nwalk)
        self._check(self.L.champ_gpu_dmc_weights(self.ptr, _p(wt), C.byref(n)), "dmc_weights")
        k = n.value
        return wt[:k], self.get_walkers()[:k], self.scalars()["eold"][:k]

    # ---------------------------------------------------------------- derivatives
    def deriv_count(self, kinds: int) -> int:
        n = C.c_int32()
        self._check(self.L.champ_gpu_deriv_count(self.ptr, int(kinds), C.byref(n)), "deriv_count")
        return n.value

    def deriv_psi(self, kinds: int) -> np.ndarray:
        """d ln psi / d p [nwalk][nparm] at the current walkers"""
        o = np.zeros((self.nwalk, self.deriv_count(kinds)))
        self._check(self.L.champ_gpu_deriv_psi(self.ptr, int(kinds), _p(o)), "deriv_psi")
        return o

    def set_deriv_map(self, kinds: int, m: Optional[np.ndarray]):
        """O' = O m for deriv_psi / sr_sums called with DERIV_MAPPED; m [nraw][nout]
        over the raw parameters of kinds (None removes the map)"""
        if m is None:
            self._check(self.L.champ_gpu_set_deriv_map(self.ptr, int(kinds), 0, None), "set_deriv_map")
            return
        m = np.ascontiguousarray(m, dtype=np.float64)
        if m.ndim != 2 or m.shape[0] != self.deriv_count(kinds):
            raise ValueError("set_deriv_map: m must be [nraw][nout] for the raw parameters of kinds")
        self._check(self.L.champ_gpu_set_deriv_map(self.ptr, int(kinds), m.shape[1], _p(m)), "set_deriv_map")

    def set_params(self, a4=None, b=None, c=None, cdet=None):
        """New Jastrow / determinant coefficients (system-array layouts; None keeps);
        the walkers' state is recomputed"""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (a4, b, c, cdet)]
        sy = self.system
        for a, ref in zip(arrs, (sy.a4, sy.b, sy.c, sy.cdet)):
            if a is not None and a.size != np.asarray(ref).size:
                raise ValueError("set_params: array size differs from the system's")
        self._check(self.L.champ_gpu_set_params(self.ptr, *[None if a is None else _p(a) for a in arrs]), "set_params")

    def set_orb_terms(self, terms):
        """Orbital terms (io, jo) of DERIV_ORB (0-based; phi_io += kappa phi_jo)"""
        t = np.ascontiguousarray(np.asarray(terms, dtype=np.int32).reshape(-1, 2))
        self._check(self.L.champ_gpu_set_orb_terms(self.ptr, len(t), t.ctypes.data_as(C.POINTER(C.c_int32))),
                    "set_orb_terms")

    def set_orbitals(self, coef):
        """New orbital coefficients [norb][nbasis]; the walkers' state is recomputed"""
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        if coef.shape != np.asarray(self.system.coef).shape:
            raise ValueError("set_orbitals: coef must have the system's shape")
        self._check(self.L.champ_gpu_set_orbitals(self.ptr, _p(coef)), "set_orbitals")

    def sr_sums(self, kinds: int):
        """(sum O O^T, sum O, sum E_L O, sum E_L) over the current walkers"""
        n = self.deriv_count(kinds)
        s, os_, eo = np.zeros((n, n)), np.zeros(n), np.zeros(n)
        e = C.c_double()
        self._check(self.L.champ_gpu_sr_sums(self.ptr, int(kinds), _p(s), _p(os_), _p(eo), C.byref(e)), "sr_sums")
        return s, os_, eo, e.value

    def sr_accumulate(self, kinds: int, reset: bool = False):
        """Adds the current walkers' SR sums to the accumulators on the GPU
        (champ_gpu_sr_accumulate); returns the running (wsum, esum)."""
        w, e = C.c_double(), C.c_double()
        self._check(self.L.champ_gpu_sr_accumulate(self.ptr, int(kinds), int(bool(reset)), C.byref(w), C.byref(e)),
                    "sr_accumulate")
        self._sr_nparm = self.deriv_count(kinds)
        return w.value, e.value

    def sr_solve(self, sr_tau: float = 0.02, sr_eps: float = 1e-3, sr_adiag: float = 0.01,
                 dparm_norm_min: float = 1.0):
        """The SR step from the accumulated sums, S z on the GPU
        (champ_gpu_sr_solve): returns (dparm, info)."""
        info = SrInfo()
        dparm = np.zeros(getattr(self, "_sr_nparm", 0))
        self._check(self.L.champ_gpu_sr_solve(self.ptr, float(sr_tau), float(sr_eps), float(sr_adiag),
                                              float(dparm_norm_min), _p(dparm), C.byref(info)), "sr_solve")
        return dparm, dict(adiag=info.adiag, dparm_norm=info.dparm_norm, cg_iters=info.cg_iters, retries=info.retries)

    def set_metrop6(self, deltar: float = 5.0, deltat: float = 1.0, fbias: float = 1.0):
        self._check(self.L.champ_gpu_set_metrop6(self.ptr, float(deltar), float(deltat), float(fbias)), "set_metrop6")

    def vmc(self, nblk: int, nstep: int, nblkeq: int = 0, tau: float = 0.5, imetro: int = 1):
        """Blocked VMC (vmc.f90 with acuest/finwrt estimators): returns the
        result fields and the block averages."""
        enow = np.zeros(int(nblk))
        r = VmcResult()
        self._check(self.L.champ_gpu_vmc(self.ptr, int(nblk), int(nstep), int(nblkeq), float(tau), int(imetro),
                                         _p(enow), C.byref(r)), "vmc")
        return dict(efin=r.efin, eerr=r.eerr, sigma=r.sigma, tcorr=r.tcorr, accept=r.accept, nsamples=r.nsamples,
                    nflagged=r.nflagged, enow=enow)

    def scalars(self) -> Dict[str, np.ndarray]:
        o = [np.zeros(self.nwalk) for _ in range(4)]
        self._check(self.L.champ_gpu_get_walker_scalars(self.ptr, *map(_p, o)), "get_walker_scalars")
        return dict(psido=o[0], psijo=o[1], eold=o[2], psi2o=o[3])

    def status(self) -> np.ndarray:
        s = np.zeros(self.nwalk, dtype=np.int32)
        self._check(self.L.champ_gpu_get_status(self.ptr, s.ctypes.data_as(C.POINTER(C.c_int32))), "get_status")
        return s

    def array(self, name: str) -> np.ndarray:
        which = ARR[name]
        n = self.L.champ_gpu_array_size(self.ptr, which)
        out = np.zeros(n)
        self._check(self.L.champ_gpu_get_array(self.ptr, which, _p(out)), f"get_array({name})")
        W, NE, NM, NO = self.nwalk, self.nelec, self.nmax, self.norb
        shapes = dict(slmi=(W, 2, NM, NM), vj=(W, NE, 3), vjn=(W, NE, 3), ddx=(W, NE, 3), d2dx2=(W, NE),
                      detiab=(W, 2), vpsp=(W, 2), peloc=(W,), d2j=(W,), orbn=(W, 4, NO), nxquad=(W,),
                      fso=(W, self.nelec, self.nelec))
        return out.reshape(shapes[name])

    def basis_fns(self, pos: np.ndarray) -> np.ndarray:
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((pos.shape[0], 5, self.system.nbasis))
        self._check(self.L.champ_gpu_basis_fns(self.ptr, pos.shape[0], _p(pos), _p(out)), "basis_fns")
        return out

    def basis_fns_noalloc(self, pos: np.ndarray) -> np.ndarray:
        """basis_fns staged in the library's static device scratch (no device allocation)"""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((pos.shape[0], 5, self.system.nbasis))
        self._check(self.L.champ_gpu_basis_fns_noalloc(self.ptr, pos.shape[0], _p(pos), _p(out)), "basis_fns_noalloc")
        return out

    # ---------------------------------------------------------------- timing
    def autotune(self) -> Dict[str, int]:
        """Time the candidate kernel forms on the current walkers and keep the fastest."""
        ch = (C.c_int32 * 4)()
        self._check(self.L.champ_gpu_autotune(self.ptr, ch), "autotune")
        return dict(move_gemm=ch[0], orbital_gemm=ch[1], ecp_rows_gemm=ch[2], quadrature=ch[3])

    def set_tuning(self, move_gemm=-1, orbital_gemm=-1, ecp_rows_gemm=-1, quadrature=-1):
        """Apply kernel choices without timing them (the layout autotune returns; -1: shape rule)."""
        ch = (C.c_int32 * 4)(move_gemm, orbital_gemm, ecp_rows_gemm, quadrature)
        self._check(self.L.champ_gpu_set_tuning(self.ptr, ch), "set_tuning")

    def set_graphs(self, on: bool = True):
        """metrop1 replays a captured HIP graph per sweep, or launches eagerly (default)."""
        self._check(self.L.champ_gpu_set_graphs(self.ptr, 1 if on else 0), "set_graphs")

    def enable_timing(self, on: bool = True):
        n = C.c_int32(1 if on else 0)
        self._check(self.L.champ_gpu_kernel_times(self.ptr, C.byref(n), None, None, None), "kernel_times")

    def kernel_times(self) -> Dict[str, tuple]:
        n = C.c_int32(16)
        names = C.create_string_buffer(32 * 16)
        ms = np.zeros(16)
        cnt = np.zeros(16, dtype=np.int64)
        self._check(self.L.champ_gpu_kernel_times(self.ptr, C.byref(n), names, _p(ms),
                                                  cnt.ctypes.data_as(C.POINTER(C.c_int64))), "kernel_times")
        out = {}
        for k in range(n.value):
            nm = names.raw[32 * k:32 * k + 32].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[k]), int(cnt[k]))
        return out


def metrop1_batch(engines, nsweep: int, tau: float = 0.5):
    """metrop1 over several populations at once (champ_gpu_metrop1_batch):
    each engine evolves as under its own metrop1 while the GPU overlaps them.
    Returns one stats dict per engine."""
    engines = list(engines)
    if not engines:
        raise ValueError("metrop1_batch: no engines")
    L = engines[0].L
    if any(e.L is not L for e in engines):
        raise ValueError("metrop1_batch: engines from different library builds")
    ptrs = (C.c_void_p * len(engines))(*[e.ptr.value for e in engines])
    st = (SweepStats * len(engines))()
    rc = L.champ_gpu_metrop1_batch(ptrs, len(engines), int(nsweep), float(tau), st)
    if rc != 0:
        raise RuntimeError(f"metrop1_batch failed ({rc}): {L.champ_gpu_last_error().decode()}")
    return [dict(accepted=s.accepted, proposed=s.proposed, esum=s.esum, esum2=s.esum2, nflagged=s.nflagged)
            for s in st]
