"""Parity of the FP64 HIP engine with the CPU oracle on identical walkers.

Bar: 1e-10 relative for psi ratios, drift vectors and local energies (the
plain FP64 path), checked walker by walker.  Every call goes through the
C-ABI (libchamp_gpu.so).  The *_md systems are multideterminant trial
functions (40-60 determinants, 1-4-fold excitations, reordered orbital slots);
the oracle evaluates them determinant by determinant.
"""
import numpy as np
import pytest

from champ_amd.engine import Engine
from tests.gpu_helpers import ENGINE_LIB, Pair, rel, relnorm

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _hpsi_both(pr: Pair):
    g = pr.eng.hpsi()
    o = [w.hpsi() for w in pr.orc]
    return g, o


@pytest.mark.parametrize("name,nwalk", [("small_ae", 6), ("small", 6), ("c22", 4), ("small_ae_md", 4),
                                        ("small_md", 4), ("c22_md", 3)])
def test_hpsi_parity(require_gpu, name, nwalk):
    pr = Pair(name, nwalk)
    pr.quad()
    g, o = _hpsi_both(pr)
    print(f"[hpsi {name}] energies gpu={g['energy']} orc={[d['energy'] for d in o]}", flush=True)
    for key in ("psid", "psij", "ekin", "energy"):
        ref = np.array([d[key] for d in o])
        assert rel(g[key], ref) < TOL, (key, g[key], ref)
    vj = pr.eng.array("vj")
    ddx = pr.eng.array("ddx")
    d2 = pr.eng.array("d2dx2")
    det = pr.eng.array("detiab")
    for w in range(nwalk):
        assert relnorm(vj[w], pr.orc[w].vj()) < TOL
        a, b = pr.orc[w].ddx()
        assert relnorm(ddx[w], a) < TOL
        assert relnorm(d2[w], b) < TOL
        assert rel(det[w], pr.orc[w].detiab()) < TOL
    assert not np.any(pr.eng.status())


@pytest.mark.parametrize("name,nwalk", [("small", 5), ("c22", 4), ("small_md", 4), ("c22_md", 3), ("het300", 1)])
def test_psie_drift_accept_parity(require_gpu, name, nwalk):
    """psie + compute_determinante_grad + accept/reject over every electron
    (het300: 150 electrons per spin, the row-by-row Sherman-Morrison form and
    the inverse in a global work matrix).

    The drifts come from the inverse as updated by every accepted move so far,
    on both sides with their own summation order.  On the small systems they
    agree to 1e-10; after 100-150 updates of a 150 x 150 inverse the rounding
    of the two update histories differs by ~1e-10 of the drift's norm, so
    het300's drifts are held to 1e-9 (ratios, Jastrow values and the fresh
    hpsi at the end to 1e-10)."""
    pr = Pair(name, nwalk, seed=5)
    pr.quad()
    _hpsi_both(pr)
    tol_v = TOL if max(pr.sys.nup, pr.sys.nelec - pr.sys.nup) <= 80 else 1e-9
    worst = 0.0
    rng = np.random.default_rng(7)
    for iel in range(pr.sys.nelec):
        vo = pr.eng.compute_determinante_grad(iel, 1)
        for w in range(nwalk):
            d = relnorm(vo[w], pr.orc[w].grad(iel, 1))
            worst = max(worst, d)
            assert d < tol_v, ("vold", iel, w, d)
        x = pr.eng.get_walkers()
        xnew = x[:, iel, :] + rng.normal(0, 0.4, size=(nwalk, 3))
        pd, pj = pr.eng.psie(iel, xnew)
        vn = pr.eng.compute_determinante_grad(iel, 0)
        for w in range(nwalk):
            opd, opj = pr.orc[w].psie(iel, xnew[w])
            assert rel(pd[w], opd) < TOL, ("psidn", iel, w, pd[w], opd)
            assert abs(pj[w] - opj) <= TOL * max(1.0, abs(opj)), ("psijn", iel, w)
            d = relnorm(vn[w], pr.orc[w].grad(iel, 0))
            worst = max(worst, d)
            assert d < tol_v, ("vnew", iel, w, d)
        mask = rng.random(nwalk) < 0.6
        pr.eng.accept(iel, mask)
        for w in range(nwalk):
            (pr.orc[w].accept if mask[w] else pr.orc[w].reject)(iel)
    print(f"[accept {name}] largest drift difference {worst:.2e}", flush=True)
    # the incrementally updated state must still agree after a full sweep
    xg = pr.eng.get_walkers()
    for w in range(nwalk):
        assert np.max(np.abs(xg[w] - pr.orc[w].get_x())) == 0.0
    slmi = pr.eng.array("slmi")
    o = [w.slmi() for w in pr.orc]
    for w in range(nwalk):
        assert relnorm(slmi[w], o[w]) < 1e-9
    g, od = _hpsi_both(pr)
    assert rel(g["energy"], [d["energy"] for d in od]) < TOL


def test_basis_fns_parity(require_gpu):
    import ctypes as C
    from oracle import ffi
    pr = Pair("c22", 1)
    pts = pr.sys.cent[pr.sys.iwctype[:]].mean(axis=0) + np.random.default_rng(1).normal(0, 1.5, size=(40, 3))
    g = pr.eng.basis_fns(pts)
    L = ffi.lib()
    nb = pr.sys.nbasis
    for p in range(len(pts)):
        rvec = (pts[p][None, :] - pr.sys.cent).copy()
        r = np.sqrt((rvec ** 2).sum(1))
        phi, dphi, d2 = np.zeros(nb), np.zeros(3 * nb), np.zeros(nb)
        L.orc_basis_fns_point(C.byref(pr.csys), ffi.P(rvec), ffi.P(r), 2, ffi.P(phi), ffi.P(dphi), ffi.P(d2))
        assert relnorm(g[p, 0], phi) < 1e-12
        assert relnorm(g[p, 1:4], dphi.reshape(3, nb)) < 1e-12
        assert relnorm(g[p, 4], d2) < 1e-12


@pytest.mark.parametrize("name", ["small", "small_md"])
def test_metrop1_trajectory(require_gpu, name):
    """Two metrop1 sweeps from the same seeds: the GPU population follows the
    oracle chains (same uniforms; gaussians agree to the last bits)."""
    from champ_amd import system as S
    from oracle import ffi
    pr = Pair(name, 6, seed=9)
    streams = ffi.seeded_streams(S.seed_from_string(), pr.nwalk)
    pr.eng.set_rng(streams)
    for w in range(pr.nwalk):
        pr.orc[w].set_rng(streams[w])
    _hpsi_both(pr)
    st = pr.eng.metrop1(2, tau=0.3)
    acc = 0
    for w in pr.orc:
        acc += w.sweep(0.3)
        acc += w.sweep(0.3)
    assert st["accepted"] == acc
    xg = pr.eng.get_walkers()
    sc = pr.eng.scalars()
    for w in range(pr.nwalk):
        assert np.max(np.abs(xg[w] - pr.orc[w].get_x())) < 1e-9
        assert abs(sc["eold"][w] - pr.orc[w].scalars()["eold"]) < 1e-9 * max(1.0, abs(sc["eold"][w]))
    rg = pr.eng.get_rng()
    for w in range(pr.nwalk):
        assert np.array_equal(rg[w], pr.orc[w].get_rng())


@pytest.mark.parametrize("name", ["small", "small_ae", "small_md"])
def test_metrop6_trajectory(require_gpu, name):
    """Two metrop6 sweeps (CHAMP's accelerated mover, metrop_mov1_slat.f90)
    from the same seeds, at its default parameters and at others: the GPU
    population follows the oracle chains draw for draw (equal RNG states and
    acceptance counts), through the rejection samplings of r and phi, the
    geometric rejections and the reverse-move densities."""
    from champ_amd import system as S
    from oracle import ffi
    for deltar, deltat, fbias in ((5.0, 1.0, 1.0), (3.0, 1.6, 0.7)):
        pr = Pair(name, 6, seed=9)
        streams = ffi.seeded_streams(S.seed_from_string(), pr.nwalk)
        pr.eng.set_rng(streams)
        for w in range(pr.nwalk):
            pr.orc[w].set_rng(streams[w])
        _hpsi_both(pr)
        pr.eng.set_metrop6(deltar, deltat, fbias)
        st = pr.eng.metrop6(2)
        acc = 0
        for w in pr.orc:
            acc += w.sweep6(deltar, deltat, fbias)
            acc += w.sweep6(deltar, deltat, fbias)
        print(f"[metrop6 {name} {deltar} {deltat} {fbias}] accepted {st['accepted']} of {st['proposed']}", flush=True)
        assert st["accepted"] == acc and st["nflagged"] == 0
        xg = pr.eng.get_walkers()
        sc = pr.eng.scalars()
        for w in range(pr.nwalk):
            assert np.max(np.abs(xg[w] - pr.orc[w].get_x())) < 1e-9
            assert abs(sc["eold"][w] - pr.orc[w].scalars()["eold"]) < 1e-9 * max(1.0, abs(sc["eold"][w]))
        rg = pr.eng.get_rng()
        for w in range(pr.nwalk):
            assert np.array_equal(rg[w], pr.orc[w].get_rng())


@pytest.mark.parametrize("name,nwalk", [("small", 4), ("c22_md", 3)])
def test_autotune_keeps_parity(require_gpu, name, nwalk):
    """The autotuner's kernel choices (GEMM kinds, quadrature form) leave hpsi
    at the oracle's values."""
    pr = Pair(name, nwalk, seed=6)
    pr.quad()
    ch = pr.eng.autotune()
    print(f"[autotune {name}] {ch}", flush=True)
    assert ch["move_gemm"] in (0, 1, 2, 4) and ch["orbital_gemm"] in (0, 1, 2)
    # all-electron systems have no ECP coefficient rows and no quadrature to tune
    ecp = pr.sys.nloc > 0
    assert ch["ecp_rows_gemm"] in ((0, 1, 2) if ecp else (-1,))
    assert ch["quadrature"] in ((0, 1) if ecp else (-1,))
    g, o = _hpsi_both(pr)
    for key in ("psid", "psij", "ekin", "energy"):
        assert rel(g[key], [d[key] for d in o]) < TOL, key


@pytest.mark.parametrize("name", ["small", "small_md"])
def test_metrop1_graph_replays_match_eager(require_gpu, name):
    """metrop1 through the captured sweep graph (replayed three times, with an
    eager hpsi and readbacks between replays) against the same sweeps launched
    eagerly from identical walkers and streams: bit-identical walkers."""
    from champ_amd import system as S
    sy = S.config(name)
    x0 = sy.initial_walkers(4, seed=3)
    engs = []
    for graphs in (True, False):
        e = Engine(sy, 4, lib_path=ENGINE_LIB)
        e.set_graphs(graphs)
        e.set_walkers(x0)
        e.seed_rng(S.seed_from_string())
        e.hpsi()
        engs.append(e)
    out = []
    for e in engs:
        s1 = e.metrop1(2, tau=0.3)
        h = e.hpsi()
        s2 = e.metrop1(1, tau=0.3)
        out.append((s1, s2, h, e.get_walkers(), e.scalars()))
    (a1, a2, ha, xa, sa), (b1, b2, hb, xb, sb) = out
    print(f"[graph {name}] accepted {a1['accepted']} {b1['accepted']} esum {a1['esum']} {b1['esum']}", flush=True)
    # walkers, per-walker energies and acceptance counts are bit-identical; the
    # estimator sums are accumulated across walkers by atomic adds, whose order
    # is not fixed, so they agree to rounding
    for sa_, sb_ in ((a1, b1), (a2, b2)):
        assert sa_["accepted"] == sb_["accepted"] and sa_["nflagged"] == sb_["nflagged"]
        for k in ("esum", "esum2"):
            assert abs(sa_[k] - sb_[k]) <= 1e-12 * max(1.0, abs(sb_[k])), k
    assert np.array_equal(xa, xb)
    assert np.array_equal(ha["energy"], hb["energy"])
    assert np.array_equal(sa["eold"], sb["eold"])


def test_large_spin_block_parity(require_gpu):
    """300 electrons (150 per spin, the largest molecule of the heterogeneous
    batch): the Slater matrices exceed LDS, so the inverse works in a global
    work matrix.  hpsi, the inverse itself, and one down-spin electron's psie
    and drift against the oracle."""
    pr = Pair("het300", 1, seed=5)
    pr.quad()
    g, o = _hpsi_both(pr)
    print(f"[het300] E gpu={g['energy']} orc={[d['energy'] for d in o]}", flush=True)
    for key in ("psid", "psij", "energy"):
        assert rel(g[key], [d[key] for d in o]) < TOL, key
    assert relnorm(pr.eng.array("slmi")[0], pr.orc[0].slmi()) < 1e-9
    iel = 231
    x = pr.eng.get_walkers()
    xnew = x[:, iel, :] + 0.3
    pd, pj = pr.eng.psie(iel, xnew)
    vn = pr.eng.compute_determinante_grad(iel, 0)
    opd, opj = pr.orc[0].psie(iel, xnew[0])
    assert rel(pd[0], opd) < TOL
    assert abs(pj[0] - opj) <= TOL * abs(opj)
    assert relnorm(vn[0], pr.orc[0].grad(iel, 0)) < TOL


@pytest.mark.parametrize("nwalk", [2])
def test_c160_parity(require_gpu, nwalk):
    """The 160-electron ECP system: hpsi and one electron's psie."""
    pr = Pair("c160", nwalk, seed=4)
    pr.quad()
    g, o = _hpsi_both(pr)
    print(f"[c160] E gpu={g['energy']} orc={[d['energy'] for d in o]} nxquad={pr.eng.array('nxquad')}", flush=True)
    for key in ("psid", "psij", "energy"):
        assert rel(g[key], [d[key] for d in o]) < TOL, key
    iel = 97
    x = pr.eng.get_walkers()
    xnew = x[:, iel, :] + 0.3
    pd, pj = pr.eng.psie(iel, xnew)
    vn = pr.eng.compute_determinante_grad(iel, 0)
    for w in range(nwalk):
        opd, opj = pr.orc[w].psie(iel, xnew[w])
        assert rel(pd[w], opd) < TOL
        assert abs(pj[w] - opj) <= TOL * abs(opj)
        assert relnorm(vn[w], pr.orc[w].grad(iel, 0)) < TOL


def test_init_seeds_default_streams(require_gpu):
    """A new population's walker streams are those of CHAMP's default seed
    string, as seed_rng(default) gives them (no all-zero xoshiro state)."""
    from champ_amd import system as S
    e = Engine(S.config("small"), 3, lib_path=ENGINE_LIB)
    r0 = e.get_rng()
    assert np.all(r0.reshape(3, 4).any(axis=1))
    e.seed_rng(S.seed_from_string())
    assert np.array_equal(r0, e.get_rng())
