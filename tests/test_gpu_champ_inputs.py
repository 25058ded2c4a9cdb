"""The engine on two of the reference's own CI cases, read from CHAMP input
files (tests/golden/champ_inputs): H2 with a numerical STO basis, all-electron,
one determinant (tests/CI_test/VMC-H2), and butadiene with BFD
pseudopotentials, a pVQZ numerical basis up to f functions and 20
determinants from 12 CSFs (tests/CI_test/VMC-C4H6-ci44_pVQZ_20_dets_12_csf_noopt).

* parity with the oracle on the same walkers (1e-10, FP64 path);
* the blocked VMC energy (champ_gpu_vmc) of both paths against the energy the
  reference's CI asserts, with its rule |E - E_ref| <= 2 sqrt(err^2 + err_ref^2)
  (tools/compare_value.py:14-28);
* the mixed path's VMC energy within one standard error of the FP64 run's on
  the same walkers and streams.
"""
import numpy as np
import pytest

from champ_amd import system as S
from champ_amd.engine import PREC_FP64, PREC_MIXED, Engine
from oracle import ffi
from tests.gpu_helpers import ENGINE_LIB, rel, relnorm
from tests.test_champ_io import load_case, reference_energy

pytestmark = pytest.mark.gpu

TOL = 1e-10
_CACHE = {}


def _case(name):
    if name not in _CACHE:
        _CACHE[name] = load_case(name)
    return _CACHE[name]


@pytest.mark.parametrize("name", ["H2", "C4H6_ci44"])
def test_ci_case_parity(require_gpu, name):
    sy = _case(name)
    cs = sy.as_ctypes()
    nw = 3
    x = sy.initial_walkers(nw, seed=7)
    eng = Engine(sy, nw, lib_path=ENGINE_LIB)
    eng.set_walkers(x)
    orc = [ffi.walker(sy, cs) for _ in range(nw)]
    for w in range(nw):
        orc[w].set_x(x[w])
        if sy.nloc:
            orc[w].set_quad(sy.xq0, sy.yq0, sy.zq0)
    g = eng.hpsi()
    o = [w.hpsi() for w in orc]
    print(f"[ci {name}] E gpu {g['energy']} oracle {[d['energy'] for d in o]}", flush=True)
    for key in ("psid", "psij", "ekin", "energy"):
        assert rel(g[key], [d[key] for d in o]) < TOL, key
    rng = np.random.default_rng(2)
    for iel in range(sy.nelec):
        vo = eng.compute_determinante_grad(iel, 1)
        xnew = eng.get_walkers()[:, iel, :] + rng.normal(0, 0.3, size=(nw, 3))
        pd, pj = eng.psie(iel, xnew)
        vn = eng.compute_determinante_grad(iel, 0)
        acc = rng.random(nw) < 0.6
        for w in range(nw):
            assert relnorm(vo[w], orc[w].grad(iel, 1)) < TOL
            opd, opj = orc[w].psie(iel, xnew[w])
            assert rel(pd[w], opd) < TOL and abs(pj[w] - opj) <= TOL * max(1.0, abs(opj))
            assert relnorm(vn[w], orc[w].grad(iel, 0)) < TOL
            (orc[w].accept if acc[w] else orc[w].reject)(iel)
        eng.accept(iel, acc)
    g = eng.hpsi()
    assert rel(g["energy"], [w.hpsi()["energy"] for w in orc]) < TOL


# walkers, equilibration blocks, blocks, sweeps per block, tau (metrop1)
VMC_RUN = {"H2": (4096, 5, 40, 20, 0.5), "C4H6_ci44": (2048, 5, 40, 10, 0.2)}


@pytest.mark.parametrize("name", ["H2", "C4H6_ci44"])
def test_ci_case_vmc_energy(require_gpu, name):
    sy = _case(name)
    W, neq, nblk, nstep, tau = VMC_RUN[name]
    eref, err_ref = reference_energy(name)
    res = {}
    for prec in (PREC_FP64, PREC_MIXED):
        e = Engine(sy, W, precision=prec, lib_path=ENGINE_LIB)
        e.set_walkers(sy.initial_walkers(W, seed=3))
        e.seed_rng(S.seed_from_string())
        e.hpsi()
        r = e.vmc(nblk, nstep, nblkeq=neq, tau=tau)
        res[prec] = r
        print(f"[vmc {name} {'fp64' if prec == PREC_FP64 else 'mixed'}] E = {r['efin']:.6f} +- {r['eerr']:.6f} "
              f"(sigma {r['sigma']:.4f}, Tcorr {r['tcorr']:.2f}, acceptance {r['accept']:.3f}); "
              f"reference {eref} +- {err_ref}", flush=True)
        assert r["nflagged"] == 0 and r["nsamples"] == W * nblk * nstep
        assert abs(r["efin"] - eref) <= 2.0 * np.hypot(r["eerr"], err_ref)
        e.close()
    d = abs(res[PREC_MIXED]["efin"] - res[PREC_FP64]["efin"])
    print(f"[vmc {name}] mixed - fp64 = {d:.2e} (one standard error {res[PREC_FP64]['eerr']:.2e})", flush=True)
    assert d <= res[PREC_FP64]["eerr"]


@pytest.mark.parametrize("name", ["H2", "C4H6_ci44"])
def test_ci_case_vmc_energy_metrop6(require_gpu, name):
    """The same CI cases sampled with CHAMP's default mover, metrop6 at its
    default deltar 5, deltat 1, fbias 1 (parser.f90:398-404): the mover the
    reference energies were produced with.  FP64 and mixed paths."""
    sy = _case(name)
    W, neq, nblk, nstep, _ = VMC_RUN[name]
    eref, err_ref = reference_energy(name)
    res = {}
    for prec in (PREC_FP64, PREC_MIXED):
        e = Engine(sy, W, precision=prec, lib_path=ENGINE_LIB)
        e.set_walkers(sy.initial_walkers(W, seed=3))
        e.seed_rng(S.seed_from_string())
        e.hpsi()
        r = e.vmc(nblk, nstep, nblkeq=neq, imetro=6)
        res[prec] = r
        print(f"[vmc6 {name} {'fp64' if prec == PREC_FP64 else 'mixed'}] E = {r['efin']:.6f} +- {r['eerr']:.6f} "
              f"(sigma {r['sigma']:.4f}, Tcorr {r['tcorr']:.2f}, acceptance {r['accept']:.3f}); "
              f"reference {eref} +- {err_ref}", flush=True)
        assert r["nflagged"] == 0 and r["nsamples"] == W * nblk * nstep
        assert abs(r["efin"] - eref) <= 2.0 * np.hypot(r["eerr"], err_ref)
        e.close()
    d = abs(res[PREC_MIXED]["efin"] - res[PREC_FP64]["efin"])
    print(f"[vmc6 {name}] mixed - fp64 = {d:.2e} (one standard error {res[PREC_FP64]['eerr']:.2e})", flush=True)
    assert d <= res[PREC_FP64]["eerr"]


def test_vmc_block_estimators(require_gpu):
    """champ_gpu_vmc's estimators from its own block averages: efin is their
    mean (equal block weights) and eerr = sqrt(var_blocks / nblk), the
    finwrt.f90:182-187 form."""
    sy = S.config("small")
    e = Engine(sy, 64, lib_path=ENGINE_LIB)
    e.set_walkers(sy.initial_walkers(64, seed=4))
    e.seed_rng(S.seed_from_string())
    e.hpsi()
    r = e.vmc(12, 3, nblkeq=1, tau=0.3)
    b = r["enow"]
    assert abs(r["efin"] - b.mean()) <= 1e-12 * max(1.0, abs(b.mean()))
    assert abs(r["eerr"] - np.sqrt(np.mean(b * b) - b.mean() ** 2) / np.sqrt(len(b))) <= 1e-9 * r["eerr"]
    r6 = e.vmc(4, 3, nblkeq=1, imetro=6)
    assert r6["nsamples"] == 64 * 4 * 3 and 0.0 < r6["accept"] < 1.0
