"""GPU tests at the sizes of the named workloads, and of the paths the small
parity cases do not reach.

* hpsi at 4096 walkers (160 electrons, ECP): the walkers on both sides of every
  boundary of the hpsi walker chunks against the oracle;
* the mixed path: the FP64 recompute of the Slater inverse (the one metrop1
  runs every 80 moves) against the oracle after FP32-fed moves, a 160-electron
  metrop1 sweep that crosses it, and the 22-electron mixed workload at 4096
  walkers against the oracle before and after sampling;
* the heterogeneous batch at its full walker counts (batched == separate) and
  each of its molecules against the oracle;
* the 5000-determinant system against the oracle;
* the FP32 per-move contraction against the FP64 one;
* a walker with a singular Slater matrix flagged without touching the others;
* dump and resume of positions and streams (dumper.f90:51-103);
* the Slater inverse at 138 and 140 electrons per spin (the LDS limit).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from champ_amd import system as S
from champ_amd.engine import PREC_FP64, PREC_MIXED, Engine, metrop1_batch
from oracle import ffi
from tests.gpu_helpers import ENGINE_LIB, rel, relnorm

pytestmark = pytest.mark.gpu

TOL = 1e-10
POOL = 8  # oracle walkers evaluated in parallel (ctypes releases the GIL)


def _rotated_quadrature(sy, csys, nwalk):
    """One rotation per walker from the default streams, as Pair.quad does."""
    import ctypes as C
    L = ffi.lib()
    streams = ffi.seeded_streams(S.seed_from_string(), nwalk)
    nq = len(sy.wq)
    X = np.zeros((nwalk, nq)); Y = np.zeros_like(X); Z = np.zeros_like(X)
    for w in range(nwalk):
        st = streams[w].copy()
        L.orc_rotqua(C.byref(csys), ffi.U(st), ffi.P(X[w]), ffi.P(Y[w]), ffi.P(Z[w]))
    return X, Y, Z


def _oracle_hpsi(sy, csys, x, quad=None):
    w = ffi.walker(sy, csys)
    w.set_x(x)
    if quad is not None:
        w.set_quad(*quad)
    return w, w.hpsi()


def _oracle_many(sy, csys, xs, quads=None):
    with ThreadPoolExecutor(POOL) as ex:
        futs = [ex.submit(_oracle_hpsi, sy, csys, xs[k], None if quads is None else quads[k]) for k in range(len(xs))]
        return [f.result() for f in futs]


def _hpsi_chunk(sy, nwalk):
    """Walkers per hpsi chunk: the all-electron orbital scratch is held to 3e9 bytes."""
    return int(max(1, min(nwalk, 3.0e9 / (sy.nelec * 5 * sy.nbasis * 8))))


# ------------------------------------------------------------ 4096 walkers
def test_hpsi_chunk_boundaries_c160(require_gpu):
    """hpsi of 4096 walkers (160 e, 600 AOs, ECP) runs in walker chunks; the
    walkers on both sides of every chunk boundary match the oracle."""
    sy = S.config("c160")
    cs = sy.as_ctypes()
    W = 4096
    x = sy.initial_walkers(W, seed=2)
    X, Y, Z = _rotated_quadrature(sy, cs, W)
    eng = Engine(sy, W, lib_path=ENGINE_LIB)
    eng.set_walkers(x)
    eng.set_quadrature(X, Y, Z)
    g = eng.hpsi()
    ch = _hpsi_chunk(sy, W)
    idx = sorted({0, W - 1} | {v for k in range(1, W // ch + 1) for v in (k * ch - 1, k * ch) if v < W})
    print(f"[chunks c160] chunk {ch} walkers; checking {idx}", flush=True)
    res = _oracle_many(sy, cs, [x[w] for w in idx], [(X[w], Y[w], Z[w]) for w in idx])
    for k, w in enumerate(idx):
        o = res[k][1]
        for key in ("psid", "psij", "ekin", "energy"):
            assert rel(g[key][w], o[key]) < TOL, (w, key, g[key][w], o[key])
    assert not np.any(eng.status())


# ------------------------------------------------------------ mixed path
# Mixed path against the FP64 oracle on identical configurations: |dE_L| /
# max(1, |E_L|), and |d ln psi| per electron (the FP32 Jastrow terms add
# rounding per pair; measured 2.4e-5 at 22 and 1.6e-4 at 256 electrons)
MIXED_TOL = dict(dlnpsi_per_e=2e-6, de_rel=2e-5)


def _mixed_ok(sy, der, dl):
    return der.max() < MIXED_TOL["de_rel"] and dl.max() < MIXED_TOL["dlnpsi_per_e"] * sy.nelec


def _mixed_errors(g, orc):
    e = np.array([o["energy"] for o in orc])
    lp = np.array([np.log(abs(o["psid"])) + o["psij"] for o in orc])
    de = np.abs(np.asarray(g["energy"]) - e)
    dl = np.abs(np.log(np.abs(np.asarray(g["psid"]))) + np.asarray(g["psij"]) - lp)
    return de, de / np.maximum(1.0, np.abs(e)), dl


def test_mixed_refresh_matches_oracle_c160(require_gpu):
    """On the mixed path the moves feed FP32 orbitals into the Sherman-Morrison
    updates; champ_gpu_determinant (the FP64 recompute metrop1 runs every 80
    moves) must give the FP64 inverse, determinants and psi_D at the walkers'
    current positions (determinant.f90:41-76)."""
    sy = S.config("c160")
    cs = sy.as_ctypes()
    nw = 2
    x0 = sy.initial_walkers(nw, seed=8)
    e32 = Engine(sy, nw, precision=PREC_MIXED, lib_path=ENGINE_LIB)
    e32.set_walkers(x0)
    e32.hpsi()
    rng = np.random.default_rng(3)
    for iel in range(90):
        x = e32.get_walkers()
        e32.psie(iel, x[:, iel, :] + rng.normal(0, 0.3, size=(nw, 3)))
        e32.accept(iel, rng.random(nw) < 0.7)
    pd = e32.determinant()
    x = e32.get_walkers()
    slmi, det = e32.array("slmi"), e32.array("detiab")
    res = _oracle_many(sy, cs, list(x))
    for w in range(nw):
        ow, o = res[w]
        assert rel(pd[w], o["psid"]) < TOL, (w, pd[w], o["psid"])
        assert rel(det[w], ow.detiab()) < TOL
        assert relnorm(slmi[w], ow.slmi()) < 1e-9
    print(f"[mixed refresh c160] psid {pd} oracle {[r[1]['psid'] for r in res]}", flush=True)


def test_mixed_sweep_crosses_refresh_c160(require_gpu):
    """One metrop1 sweep of 160 electrons on the mixed path (the inverse
    recomputed after move 80) beside the FP64 path from the same walkers and
    streams: same decisions, positions within FP32 drift rounding; the mixed
    local energies at the sampled positions against the oracle."""
    sy = S.config("c160")
    cs = sy.as_ctypes()
    nw = 4
    x0 = sy.initial_walkers(nw, seed=9)
    engs = []
    for prec in (PREC_FP64, PREC_MIXED):
        e = Engine(sy, nw, precision=prec, lib_path=ENGINE_LIB)
        e.set_walkers(x0)
        e.seed_rng(S.seed_from_string())
        e.hpsi()
        engs.append(e)
    e64, e32 = engs
    e32.enable_timing(True)
    s32 = e32.metrop1(1, tau=0.2)
    kt = e32.kernel_times()
    e32.enable_timing(False)
    s64 = e64.metrop1(1, tau=0.2)
    # the mixed sweep recomputes the inverse once beside hpsi's own recompute
    assert kt["hpsi_slater_lu"][1] == 2 * kt["hpsi_jastrow"][1], kt
    assert s32["accepted"] == s64["accepted"], (s32, s64)
    x32, x64 = e32.get_walkers(), e64.get_walkers()
    dx = float(np.max(np.abs(x32 - x64)))
    print(f"[mixed sweep c160] accepted {s32['accepted']} of {s32['proposed']}; max |dx| {dx:.2e}", flush=True)
    # the FP32 Jastrow gradient in the drift moves proposals by ~1e-6 of the
    # drift step; 160 moves leave the positions within 1e-3 bohr (2.4e-4 measured)
    assert dx < 1e-3
    sc = e32.scalars()
    g = dict(energy=sc["eold"], psid=sc["psido"], psij=sc["psijo"])
    # the sweep's hpsi used the rotation each walker's stream drew first in
    # the sweep, i.e. rotqua applied to the freshly seeded streams
    X, Y, Z = _rotated_quadrature(sy, cs, nw)
    res = _oracle_many(sy, cs, list(x32), [(X[w], Y[w], Z[w]) for w in range(nw)])
    de, der, dl = _mixed_errors(g, [r[1] for r in res])
    print(f"[mixed sweep c160] |dE_L| {de}, rel {der}; |dln psi| {dl}", flush=True)
    assert _mixed_ok(sy, der, dl)
    assert s32["nflagged"] == 0


def test_config3_c22_mixed_4096_against_oracle(require_gpu):
    """The mixed-precision workload on the 22-electron system at 4096 walkers:
    per-walker log psi and local energy against the FP64 oracle on identical
    configurations, at the initial walkers and after two sampled sweeps."""
    sy = S.config("c22")
    cs = sy.as_ctypes()
    W = 4096
    e32 = Engine(sy, W, precision=PREC_MIXED, lib_path=ENGINE_LIB)
    e32.set_walkers(sy.initial_walkers(W, seed=4))
    e32.seed_rng(S.seed_from_string())
    pick = np.arange(0, W, 64)
    for stage in ("initial", "sampled"):
        if stage == "sampled":
            st = e32.metrop1(2, tau=0.2)
            assert st["nflagged"] == 0
            sc = e32.scalars()
            g = dict(energy=sc["eold"], psid=sc["psido"], psij=sc["psijo"])
        else:
            g = e32.hpsi()
        x = e32.get_walkers()
        res = _oracle_many(sy, cs, [x[w] for w in pick])
        de, der, dl = _mixed_errors({k: np.asarray(v)[pick] for k, v in g.items()}, [r[1] for r in res])
        print(f"[config3 c22 {stage}] |dE_L| max {de.max():.2e} Ha p99 {np.percentile(de, 99):.2e} mean "
              f"{de.mean():.2e}; rel max {der.max():.2e}; |dln psi| max {dl.max():.2e}", flush=True)
        assert _mixed_ok(sy, der, dl)


# ------------------------------------------------------------ batch
def test_config4_hetero_batch_full_size(require_gpu):
    """The eight molecules (20-300 electrons) at their full walker counts
    (4096 ... 512), mixed precision: one batched sweep equals each
    population's own sweep bit for bit."""
    mols = S.hetero_batch()

    def make():
        out = []
        for k, (sy, W) in enumerate(mols):
            e = Engine(sy, W, precision=PREC_MIXED, lib_path=ENGINE_LIB)
            e.set_walkers(sy.initial_walkers(W, seed=31 + k))
            e.seed_rng(S.seed_from_string())
            e.hpsi()
            out.append(e)
        return out

    batch = make()
    sb = metrop1_batch(batch, 1, tau=0.3)
    xb = [e.get_walkers() for e in batch]
    eb = [e.scalars()["eold"] for e in batch]
    for e in batch:
        e.close()
    single = make()
    for m, e in enumerate(single):
        s1 = e.metrop1(1, tau=0.3)
        name = mols[m][0].name
        print(f"[config4 {name} W={mols[m][1]}] accepted {sb[m]['accepted']} {s1['accepted']}", flush=True)
        assert sb[m]["accepted"] == s1["accepted"] and sb[m]["nflagged"] == s1["nflagged"] == 0, name
        assert np.array_equal(xb[m], e.get_walkers()), name
        assert np.array_equal(eb[m], e.scalars()["eold"]), name
        e.close()


@pytest.mark.parametrize("name", [f"het{ne}" for ne, _, _, _ in S.HETERO])
def test_config4_molecules_against_oracle(require_gpu, name):
    """Each molecule of the batch on the FP64 path: hpsi, one electron's psie
    and drift against the oracle on two walkers."""
    sy = S.config(name)
    cs = sy.as_ctypes()
    nw = 2
    x = sy.initial_walkers(nw, seed=5)
    X, Y, Z = _rotated_quadrature(sy, cs, nw)
    eng = Engine(sy, nw, lib_path=ENGINE_LIB)
    eng.set_walkers(x)
    eng.set_quadrature(X, Y, Z)
    g = eng.hpsi()
    res = _oracle_many(sy, cs, list(x), [(X[w], Y[w], Z[w]) for w in range(nw)])
    for w in range(nw):
        for key in ("psid", "psij", "energy"):
            assert rel(g[key][w], res[w][1][key]) < TOL, (name, w, key)
    iel = sy.nelec - 3
    xnew = x[:, iel, :] + 0.25
    pd, pj = eng.psie(iel, xnew)
    vn = eng.compute_determinante_grad(iel, 0)
    for w in range(nw):
        ow = res[w][0]
        opd, opj = ow.psie(iel, xnew[w])
        assert rel(pd[w], opd) < TOL and abs(pj[w] - opj) <= TOL * max(1.0, abs(opj))
        assert relnorm(vn[w], ow.grad(iel, 0)) < TOL


# ------------------------------------------------------------ multideterminant
def test_config5_md256_against_oracle(require_gpu):
    """256 electrons, 960 AOs, 5000 determinants (1-4-fold excitations within
    40 virtuals): hpsi and one electron's psie and drift against the oracle,
    which evaluates every determinant on its own; the mixed path's local
    energy on the same walker against the same oracle value."""
    sy = S.config("md256")
    cs = sy.as_ctypes()
    x = sy.initial_walkers(1, seed=6)
    X, Y, Z = _rotated_quadrature(sy, cs, 1)
    iel = 200
    xnew = x[:, iel, :] + 0.2
    with ThreadPoolExecutor(1) as ex:  # the oracle (about a minute) runs beside the GPU work
        fut = ex.submit(lambda: _psie_after_hpsi(sy, cs, x[0], (X[0], Y[0], Z[0]), iel, xnew[0]))
        engs = {}
        out = {}
        for prec in (PREC_FP64, PREC_MIXED):
            e = Engine(sy, 1, precision=prec, lib_path=ENGINE_LIB)
            e.set_walkers(x)
            e.set_quadrature(X, Y, Z)
            g = e.hpsi()
            pd, pj = e.psie(iel, xnew)
            out[prec] = (g, pd[0], pj[0], e.compute_determinante_grad(iel, 0)[0])
            engs[prec] = e
        o, opd, opj, ovn = fut.result()
    g, pd, pj, vn = out[PREC_FP64]
    print(f"[config5 md256] E gpu {g['energy'][0]} oracle {o['energy']}", flush=True)
    for key in ("psid", "psij", "energy"):
        assert rel(g[key][0], o[key]) < TOL, key
    assert rel(pd, opd) < TOL and abs(pj - opj) <= TOL * max(1.0, abs(opj))
    assert relnorm(vn, ovn) < TOL
    gm = out[PREC_MIXED][0]
    de, der, dl = _mixed_errors(gm, [o])
    print(f"[config5 md256 mixed] |dE_L| {de[0]:.3e} Ha (rel {der[0]:.2e}); |dln psi| {dl[0]:.2e}", flush=True)
    assert _mixed_ok(sy, der, dl)


def _psie_after_hpsi(sy, cs, x, quad, iel, xnew):
    ow, o = _oracle_hpsi(sy, cs, x, quad)
    opd, opj = ow.psie(iel, xnew)
    return o, opd, opj, ow.grad(iel, 0)


# ------------------------------------------------------------ FP32 contraction
@pytest.mark.parametrize("name", ["c22", "c160"])
def test_fp32_move_contraction(require_gpu, name):
    """The mixed path's FP32 per-move contraction (GEMM kind 3; 48 orbitals per
    wavefront, so 11 and 80 orbitals both end in a masked tile) against the
    FP64 contraction of the same FP32 AO values, and the ratio against the
    oracle."""
    sy = S.config(name)
    nw = 4
    x = sy.initial_walkers(nw, seed=12)
    iel = 7
    xnew = x[:, iel, :] + 0.3
    out = {}
    for kind in (0, 3):
        e = Engine(sy, nw, precision=PREC_MIXED, lib_path=ENGINE_LIB)
        e.set_walkers(x)
        e.hpsi()
        e.set_tuning(move_gemm=kind, ecp_rows_gemm=-1, quadrature=-1)
        pd, _ = e.psie(iel, xnew)
        out[kind] = (e.array("orbn"), pd)
    d = relnorm(out[3][0], out[0][0])
    r = rel(out[3][1], out[0][1])
    print(f"[fp32 gemm {name}] orbn relnorm {d:.2e}, psidn rel {r:.2e}", flush=True)
    assert d < 1e-5 and r < 1e-4
    # the orbitals at the proposed position against the FP64 oracle (the ratio
    # itself sums them with an inverse whose condition number amplifies the
    # FP32 AO rounding on these random-orbital systems).  The AOs are FP32
    # values (unit roundoff 6e-8) summed over 100-600 AOs with cancellation:
    # 5e-5 of the largest orbital value bounds what was measured (1.1e-5).
    cs = sy.as_ctypes()
    for w in range(nw):
        ow, _ = _oracle_hpsi(sy, cs, x[w])
        ow.psie(iel, xnew[w])
        assert relnorm(out[3][0][w], ow.orbn()) < 5e-5


@pytest.mark.parametrize("name,prec", [("c22", PREC_FP64), ("c160", PREC_FP64), ("c160", PREC_MIXED)])
def test_splitk_move_contraction(require_gpu, name, prec):
    """The per-move contraction split over the two halves of K (GEMM kind 4:
    the halves' partials added in a fixed order) against the one-pass tile
    form on the same AO values (the sums differ only in their association,
    1e-13 relative), and a second call giving the same bits."""
    sy = S.config(name)
    nw = 4
    x = sy.initial_walkers(nw, seed=12)
    iel = 7
    xnew = x[:, iel, :] + 0.3
    out = {}
    for kind in (0, 4, 4):
        e = Engine(sy, nw, precision=prec, lib_path=ENGINE_LIB)
        e.set_walkers(x)
        e.hpsi()
        e.set_tuning(move_gemm=kind, ecp_rows_gemm=-1, quadrature=-1)
        pd, _ = e.psie(iel, xnew)
        out.setdefault(kind, []).append((e.array("orbn"), pd))
        e.close()
    (o0, p0), = out[0]
    (o4, p4), (o4b, p4b) = out[4]
    d = relnorm(o4, o0)
    print(f"[split-K gemm {name}] orbn relnorm {d:.2e}, psidn rel {rel(p4, p0):.2e}", flush=True)
    assert d < 1e-13 and rel(p4, p0) < 1e-11
    assert np.array_equal(o4, o4b) and np.array_equal(p4, p4b)


def test_set_tuning_rejects_fp32_on_fp64_path(require_gpu):
    e = Engine(S.config("small"), 2, lib_path=ENGINE_LIB)
    with pytest.raises(RuntimeError, match="not available"):
        e.set_tuning(move_gemm=3)


# ------------------------------------------------------------ failure path
@pytest.mark.parametrize("name", ["small", "small_ae"])
def test_singular_walker_is_flagged_alone(require_gpu, name):
    """An electron beyond every basis function's cut radius gives its Slater
    matrix a zero row: CHAMP's matinv stops the run (matinv.f90:59-62); the
    engine flags that walker CHAMP_WALKER_SINGULAR and leaves every other
    walker exactly as in a population without it, through hpsi and a sweep."""
    sy = S.config(name)
    nw = 4
    x = sy.initial_walkers(nw, seed=14)
    bad = x.copy()
    bad[2, 0, :] = [1000.0, 0.0, 0.0]
    runs = []
    for xx in (bad, x):
        e = Engine(sy, nw, lib_path=ENGINE_LIB)
        e.set_walkers(xx)
        e.seed_rng(S.seed_from_string())
        g = e.hpsi()
        st0 = e.status()
        s = e.metrop1(1, tau=0.3)
        runs.append((g, st0, e.get_walkers(), e.scalars(), e.status(), s))
    (gb, stb, xb, scb, stb2, sb), (gc, stc, xc, scc, stc2, sc) = runs
    print(f"[singular {name}] status after hpsi {stb}, after sweep {stb2}", flush=True)
    assert stb[2] & 1 and not np.any(stc)
    ok = [0, 1, 3]
    for key in ("psid", "psij", "energy"):
        assert np.array_equal(gb[key][ok], gc[key][ok]), key
    assert not np.any(stb[ok]) and not np.any(stb2[ok])
    assert np.array_equal(xb[ok], xc[ok])
    assert np.array_equal(scb["eold"][ok], scc["eold"][ok])


# ------------------------------------------------------------ restart
@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_dump_and_resume(require_gpu, prec):
    """Positions and per-walker streams dumped after two sweeps and loaded into
    a new population (the restart of dumper.f90:51-103 / startr) continue
    exactly as the original population does."""
    sy = S.config("small")
    nw = 5
    a = Engine(sy, nw, precision=prec, lib_path=ENGINE_LIB)
    a.set_walkers(sy.initial_walkers(nw, seed=15))
    a.seed_rng(S.seed_from_string())
    a.hpsi()
    a.metrop1(2, tau=0.3)
    xd, rd = a.get_walkers(), a.get_rng()
    sa = a.metrop1(2, tau=0.3)
    b = Engine(sy, nw, precision=prec, lib_path=ENGINE_LIB)
    b.set_walkers(xd)
    b.set_rng(rd)
    b.hpsi()
    sb = b.metrop1(2, tau=0.3)
    assert sa["accepted"] == sb["accepted"]
    assert np.array_equal(a.get_walkers(), b.get_walkers())
    assert np.array_equal(a.scalars()["eold"], b.scalars()["eold"])
    assert np.array_equal(a.get_rng(), b.get_rng())


def test_set_rng_rejects_zero_state(require_gpu):
    e = Engine(S.config("small"), 2, lib_path=ENGINE_LIB)
    st = e.get_rng()
    st[1] = 0
    with pytest.raises(RuntimeError, match="all-zero"):
        e.set_rng(st)


# ------------------------------------------------------------ LDS limit
@pytest.mark.parametrize("name", ["ae276", "ae280"])
def test_slater_inverse_at_the_lds_limit(require_gpu, name):
    """138 electrons per spin: the Slater work matrix and the inverse kernel's
    pivot arrays just fit one block's LDS; 140: they do not, and the global
    work matrix is used.  hpsi and the inverse against the oracle."""
    sy = S.config(name)
    cs = sy.as_ctypes()
    x = sy.initial_walkers(1, seed=16)
    e = Engine(sy, 1, lib_path=ENGINE_LIB)
    e.set_walkers(x)
    g = e.hpsi()
    ow, o = _oracle_hpsi(sy, cs, x[0])
    for key in ("psid", "psij", "energy"):
        assert rel(g[key][0], o[key]) < TOL, key
    assert relnorm(e.array("slmi")[0], ow.slmi()) < 1e-9
