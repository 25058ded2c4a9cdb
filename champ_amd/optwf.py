"""CHAMP's optimisation parameters over the engine's raw derivatives.

The engine returns d ln psi / d p for the raw coefficient arrays of a system
(a4, b, c, cdet; ``Engine.deriv_psi``).  CHAMP optimises a selection of them:
the Jastrow coefficients the jastrow_der file lists (iwjasa / iwjasb / iwjasc,
read_data_parser.f90:1521-1700), with the e-e-n coefficients fixed by the cusp
conditions following the free ones (cuspinit4 / cuspexact4, and the chain
rule of deriv_jastrow4.f90:394-460), and the determinant (or CSF)
coefficients but the first (or the largest CSF) (optci.f90:63-80,
parser.f90:1096-1113).  :class:`ChampParams` holds that selection as a matrix
M [nraw][nparm] (O_champ = O_raw M), which the engine applies on the GPU
(``Engine.set_deriv_map``), and applies a step the way compute_jastrow /
compute_ci do (optwf_handle_wf.f90:1050-1104, 1214-1284).  Orbital
excitations (io, jo) of optorb come last (phi_io += kappa phi_jo, the update
of compute_lcao, optwf_handle_wf.f90:1106-1150, with orbitals_ortho off)."""
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import champ_io as io
from . import system as S
from .engine import DERIV_CI, DERIV_JAS_EE, DERIV_JAS_EEN, DERIV_JAS_EN, DERIV_ORB


def cusp_jacobian(nordc: int) -> np.ndarray:
    """P [nterms][nterms]: column j is the full e-e-n coefficient vector when
    coefficient j is 1, the other free ones 0, and the cusp-fixed ones follow
    (cuspexact4 is linear, so c_full = P c_free for any free c)."""
    nt = S.nterms4(nordc)
    p = np.zeros((nt, nt))
    for j in range(nt):
        e = np.zeros((1, nt))
        e[0, j] = 1.0
        p[:, j] = S.cusp_fix_c(e, nordc)[0]
    return p


def cusp_fixed_terms(nordc: int) -> List[int]:
    """0-based indices of the e-e-n coefficients the cusp conditions fix
    (iwc4 of cuspinit4.f90:3-70)."""
    nt = S.nterms4(nordc)
    p = cusp_jacobian(nordc)
    return [t for t in range(nt) if p[t, t] == 0.0]


def optorb_define(iworbd, nup: int, irrep, ncore: int = 0, nadorb: int = -1, no_active: int = 0,
                  iwmix_virt=None) -> np.ndarray:
    """CHAMP's list of orbital excitations phi_io += kappa phi_jo (optorb_define,
    optorb.f90:838-1060), 0-based [nterm][2] (io, jo) in its order: io over
    the orbitals the determinants use (ncore..ndetorb-1), jo over those and
    the first ``nadorb`` external orbitals (nextorb; -1 or more than there are
    = all of them, optorb.f90:888), keeping a pair when io is occupied in some
    determinant, io != jo, both have the same irrep, not both are doubly
    occupied in every determinant, not (no_active and io partly occupied and
    jo occupied somewhere), ``iwmix_virt`` [norb][norb] (None = every pair)
    allows it, and some determinant has io occupied and jo empty in one spin.
    ``irrep`` has one entry per orbital of the orbital file."""
    iw = np.asarray(iworbd, dtype=np.int64)
    irrep = np.asarray(irrep)
    ndet, nelec = iw.shape
    norb_tot = len(irrep)
    ndetorb = int(iw.max()) + 1
    if ndetorb > norb_tot:
        raise ValueError(f"determinants use {ndetorb} orbitals, the symmetry list has {norb_tot}")
    next_max = norb_tot - ndetorb
    if nadorb < 0 or nadorb > next_max:
        nadorb = next_max
    nall = ndetorb + nadorb
    # per orbital: occupied in some determinant; doubly occupied in every one
    cnt = np.zeros((ndet, nall), dtype=np.int64)
    for k in range(ndet):
        np.add.at(cnt[k], iw[k], 1)
    used = np.zeros(nall, dtype=bool)
    used[:ndetorb] = (cnt[:, :ndetorb] > 0).any(axis=0)
    not_double = np.ones(nall, dtype=bool)
    not_double[:ndetorb] = (cnt[:, :ndetorb] != 2).any(axis=0)
    # occupation per spin block: occ[s][k][o]
    occ = np.zeros((2, ndet, nall), dtype=bool)
    for k in range(ndet):
        occ[0, k, iw[k, :nup]] = True
        occ[1, k, iw[k, nup:]] = True
    terms = []
    for io in range(ncore, ndetorb):
        if not used[io]:
            continue
        for jo in range(ncore, nall):
            if io == jo or irrep[io] != irrep[jo]:
                continue
            if not not_double[io] and not not_double[jo]:
                continue
            if no_active and not_double[io] and used[jo]:
                continue
            if iwmix_virt is not None and not iwmix_virt[io][jo]:
                continue
            if not (occ[:, :, io] & ~occ[:, :, jo]).any():
                continue
            terms.append((io, jo))
    return np.array(terms, dtype=np.int32).reshape(-1, 2)


@dataclass
class ChampParams:
    kinds: int
    m: np.ndarray                # [nraw][nparm]
    names: List[str]
    iwjasa: List[List[int]]
    iwjasb: List[List[int]]
    iwjasc: List[List[int]]
    ci: bool
    csf: Optional[Tuple[np.ndarray, list]] = None
    maxcsf: int = 0
    orb_terms: Optional[np.ndarray] = None  # [nterm][2] (io, jo), 0-based

    @property
    def nparm(self) -> int:
        return self.m.shape[1]

    @staticmethod
    def build(sy: S.System, der: Optional[dict] = None, ci: bool = False, csf=None,
              orb_terms=None) -> "ChampParams":
        """The parameter list of optwf with ioptjas = 1 (the lists of ``der``,
        from :func:`champ_io.read_jastrow_der`; None = no Jastrow
        parameters), ioptci = ``ci`` (``csf``: :func:`champ_io.read_csf`
        of the determinant file, None for plain determinants) and ioptorb
        with the excitations ``orb_terms`` ([nterm][2] (io, jo), 0-based;
        the engine needs the same list through Engine.set_orb_terms)."""
        nct, nsb = sy.nctype, sy.nspin2b
        na, nb = sy.norda + 1, sy.nordb + 1
        nt = S.nterms4(sy.nordc) if sy.nordc > 1 else 0
        ndet = len(sy.cdet)
        iwjasa = der["iwjasa"] if der else [[] for _ in range(nct)]
        iwjasb = der["iwjasb"] if der else [[] for _ in range(nsb)]
        iwjasc = der["iwjasc"] if der else [[] for _ in range(nct)]
        kinds = 0
        if any(iwjasa):
            kinds |= DERIV_JAS_EN
        if any(iwjasb):
            kinds |= DERIV_JAS_EE
        if any(iwjasc):
            kinds |= DERIV_JAS_EEN
        if ci and ndet > 1:
            kinds |= DERIV_CI
        terms = None if orb_terms is None else np.asarray(orb_terms, dtype=np.int32).reshape(-1, 2)
        if terms is not None and len(terms):
            kinds |= DERIV_ORB
        off, o = {}, 0
        for k, size in ((DERIV_JAS_EN, nct * na), (DERIV_JAS_EE, nsb * nb), (DERIV_JAS_EEN, nct * nt),
                        (DERIV_CI, ndet if ndet > 1 else 0), (DERIV_ORB, 0 if terms is None else len(terms))):
            if kinds & k:
                off[k] = o
                o += size
        nraw = o
        cols, names = [], []  # a column is a raw index (a unit vector) or a dense vector

        def unit(i):
            return int(i)

        for it in range(nct):
            for iw in iwjasa[it]:
                if not 1 <= iw <= na:
                    raise ValueError(f"iwjasa {iw} outside 1..norda+1")
                cols.append(unit(off[DERIV_JAS_EN] + it * na + iw - 1))
                names.append(f"a{iw}[{it}]")
        for isb in range(nsb):
            for iw in iwjasb[isb]:
                if not 1 <= iw <= nb:
                    raise ValueError(f"iwjasb {iw} outside 1..nordb+1")
                cols.append(unit(off[DERIV_JAS_EE] + isb * nb + iw - 1))
                names.append(f"b{iw}[{isb}]")
        if any(iwjasc):
            p = cusp_jacobian(sy.nordc)
            fixed = set(cusp_fixed_terms(sy.nordc))
            for it in range(nct):
                for iw in iwjasc[it]:
                    if not 1 <= iw <= nt or iw - 1 in fixed:
                        raise ValueError(f"iwjasc {iw}: not a free e-e-n coefficient")
                    v = np.zeros(nraw)
                    v[off[DERIV_JAS_EEN] + it * nt: off[DERIV_JAS_EEN] + (it + 1) * nt] = p[:, iw - 1]
                    cols.append(v)
                    names.append(f"c{iw}[{it}]")
        maxcsf = 0
        if kinds & DERIV_CI:
            if csf is None:
                for k in range(1, ndet):
                    cols.append(unit(off[DERIV_CI] + k))
                    names.append(f"cdet{k + 1}")
            else:
                ccsf, csf_terms = csf
                maxcsf = int(np.argmax(np.abs(ccsf)))  # first of the largest, parser.f90:1096-1109
                for icsf, t in enumerate(csf_terms):
                    if icsf == maxcsf:
                        continue
                    v = np.zeros(nraw)
                    for idet, cx in t:
                        v[off[DERIV_CI] + idet] += cx
                    cols.append(v)
                    names.append(f"csf{icsf + 1}")
        if kinds & DERIV_ORB:
            for k, (io, jo) in enumerate(terms):
                cols.append(unit(off[DERIV_ORB] + k))
                names.append(f"orb{io + 1}->{jo + 1}")
        m = np.zeros((nraw, len(cols)))
        for j, col in enumerate(cols):
            if isinstance(col, int):
                m[col, j] = 1.0
            else:
                m[:, j] = col
        return ChampParams(kinds, m, names, iwjasa, iwjasb, iwjasc, bool(kinds & DERIV_CI),
                           csf, maxcsf, terms if kinds & DERIV_ORB else None)

    def apply(self, sy: S.System, dparm: np.ndarray):
        """(a4, b, c, cdet) after the step p - dparm (compute_jastrow and
        compute_ci; the cusp-fixed e-e-n coefficients recomputed by
        cuspexact4).  ``sy`` is not modified."""
        dparm = np.asarray(dparm, dtype=np.float64)
        if dparm.shape != (self.nparm,):
            raise ValueError("apply: dparm has the wrong length")
        a4, b, c, cdet = sy.a4.copy(), sy.b.copy(), np.array(sy.c, dtype=np.float64).copy(), sy.cdet.copy()
        i = 0
        for it, lst in enumerate(self.iwjasa):
            for iw in lst:
                a4[it, iw - 1] -= dparm[i]
                i += 1
        for isb, lst in enumerate(self.iwjasb):
            for iw in lst:
                b[isb, iw - 1] -= dparm[i]
                i += 1
        if any(self.iwjasc):
            for it, lst in enumerate(self.iwjasc):
                for iw in lst:
                    c[it, iw - 1] -= dparm[i]
                    i += 1
            c = S.cusp_fix_c(c, sy.nordc)
        if self.ci:
            if self.csf is None:
                cdet[1:] -= dparm[i:i + len(cdet) - 1]
                i += len(cdet) - 1
            else:
                _, terms = self.csf
                for icsf, t in enumerate(terms):
                    if icsf == self.maxcsf:
                        continue
                    for idet, cx in t:
                        cdet[idet] -= dparm[i] * cx
                    i += 1
        if self.orb_terms is not None:
            i += len(self.orb_terms)
        assert i == self.nparm
        return a4, b, c, cdet

    def orbitals(self, sy: S.System, dparm: np.ndarray) -> np.ndarray:
        """The orbital coefficients after the step (compute_lcao: coef[io] -=
        dparm * coef[jo] with the coefficients before the step)."""
        coef = np.array(sy.coef, dtype=np.float64)
        if self.orb_terms is None:
            return coef
        dparm = np.asarray(dparm, dtype=np.float64)
        k0 = self.nparm - len(self.orb_terms)
        out = coef.copy()
        for k, (io, jo) in enumerate(self.orb_terms):
            out[io] -= dparm[k0 + k] * coef[jo]
        return out
