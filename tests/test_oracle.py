"""Pinning the CPU oracle (CPU only).

Each check is chosen so that a plausible slip (dropped term, wrong sign or
index, transposed operand) in the oracle's restatement fails it:
  * known-answer streams of CHAMP's own xoshiro C file (fixture + live build),
  * the reference unit tests for matinv and pot_nn,
  * closed forms (Gaussian contractions, sphere averages, addition theorem),
  * finite differences of ln psi for drift and kinetic energy,
  * internal identities (incremental == full recompute, SM == fresh LU),
  * cusps and asymptotes of the Jastrow factor,
  * a brute-force re-evaluation of the non-local quadrature from full psi values.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from champ_amd import system as S
from oracle import ffi

HERE = os.path.dirname(os.path.abspath(__file__))


# ----------------------------------------------------------------- RNG
def _fixture():
    with open(os.path.join(HERE, "golden", "rng_vectors.json")) as f:
        return json.load(f)


def test_rng_matches_reference_fixture():
    fx = _fixture()
    assert fx["seed_i4"] == [int(v) for v in S.seed_from_string()]
    L = ffi.lib()
    streams = ffi.seeded_streams(S.seed_from_string(), len(fx["streams"]))
    for st in fx["streams"]:
        s = streams[st["rank"]].copy()
        assert [int(v) for v in s] == [int(v) for v in st["state"]]
        nx = [int(L.orc_xoshiro_next(ffi.U(s))) for _ in range(8)]
        assert nx == [int(v) for v in st["next"]]
        s = streams[st["rank"]].copy()
        un = [L.orc_random_dp(ffi.U(s)) for _ in range(8)]
        assert un == [float(v) for v in st["uniform"]]


def test_rng_matches_reference_build():
    R = ffi.ref_rng_lib()
    if R is None:
        pytest.skip("reference RNG library not built here (needs the CHAMP tree)")
    L = ffi.lib()
    seed = S.seed_from_string()
    R.xoroshiro_set_state((C.c_uint32 * 8)(*[int(v) for v in seed]))
    mine = ffi.seeded_streams(seed, 1)[0].copy()
    for _ in range(2000):
        assert int(R.next()) == int(L.orc_xoshiro_next(ffi.U(mine)))
    R.jump()
    L.orc_xoshiro_jump(ffi.U(mine))
    for _ in range(100):
        assert R.next_uniform_double() == L.orc_random_dp(ffi.U(mine))


def test_gauss_distribution():
    L = ffi.lib()
    s = ffi.seeded_streams(S.seed_from_string(), 1)[0].copy()
    g = np.array([L.orc_gauss(ffi.U(s)) for _ in range(200000)])
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01
    # Box-Muller as written: cos(2 pi u1) * sqrt(-2 ln u2), u1 drawn first
    s = ffi.seeded_streams(S.seed_from_string(), 1)[0].copy()
    s2 = s.copy()
    u1 = L.orc_random_dp(ffi.U(s2)); u2 = L.orc_random_dp(ffi.U(s2))
    assert L.orc_gauss(ffi.U(s)) == math.cos(2 * math.pi * u1) * math.sqrt(-2 * math.log(u2))


# ------------------------------------------------------- reference unit tests
def test_matinv_reference_cases():
    """tests/unit_test/test_matinv.f90:10-39 (column-major storage)."""
    L = ffi.lib()
    a = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float64)
    det = np.zeros(1)
    assert L.orc_matinv(ffi.P(a), 3, ffi.P(det)) == 0
    assert list(a) == [1, 0, 0, 0, 1, 0, 0, 0, 1] and det[0] == 1.0
    b = np.array([1, 0, 0, 5, 1, 0, 0, 0, 1], dtype=np.float64)
    assert L.orc_matinv(ffi.P(b), 3, ffi.P(det)) == 0
    assert list(b) == [1, 0, 0, -5, 1, 0, 0, 0, 1] and det[0] == 1.0


@pytest.mark.parametrize("n", [1, 2, 3, 7, 33])
def test_matinv_random(n):
    L = ffi.lib()
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n))
    a = np.asfortranarray(A).ravel(order="F").copy()
    det = np.zeros(1)
    L.orc_matinv(ffi.P(a), n, ffi.P(det))
    Ainv = a.reshape((n, n), order="F")
    assert np.max(np.abs(Ainv @ A - np.eye(n))) < 1e-11
    assert abs(det[0] / np.linalg.det(A) - 1) < 1e-11


def test_pot_nn_water():
    """tests/unit_test/test_electron_repulsion.f90: water pot_nn = 6.98361052364638."""
    sy = S.config("small")
    sy.cent = np.array([[0.0, 0.0, 0.0], [-1.43042870599673, 0.0, -1.10715696499747],
                        [1.43042870599673, 0.0, -1.10715696499747]])
    sy.iwctype = np.array([0, 1, 1], dtype=np.int32)
    sy.znuc = np.array([6.0, 1.0])
    cs = sy.as_ctypes()
    assert abs(ffi.lib().orc_pot_nn(C.byref(cs)) - 6.98361052364638) < 1e-12


# ------------------------------------------------------------ quadrature
def _sphere_avg(a, b, c):
    if a % 2 or b % 2 or c % 2:
        return 0.0
    def df(n):
        r = 1
        while n > 1:
            r *= n
            n -= 2
        return r
    n = a + b + c
    return df(a - 1) * df(b - 1) * df(c - 1) / df(n + 1)


def test_icosahedral_rule_exact_to_degree5():
    L = ffi.lib()
    X, Y, Z, Wq = (np.zeros(86) for _ in range(4))
    L.orc_gesqua(12, ffi.P(X), ffi.P(Y), ffi.P(Z), ffi.P(Wq))
    xq, yq, zq, wq = S.icosa12()
    assert np.array_equal(X[:12], xq) and np.array_equal(Wq[:12], wq)
    assert abs(wq.sum() - 1) < 1e-15
    for a in range(6):
        for b in range(6 - a):
            for c in range(6 - a - b):
                q = np.sum(wq * X[:12] ** a * Y[:12] ** b * Z[:12] ** c)
                # pi in gesqua is 3.14159265359 (rotqua.f90), hence 1e-11 not 1e-15
                assert abs(q - _sphere_avg(a, b, c)) < 1e-11, (a, b, c)


def test_rotqua_is_a_rotation():
    sy = S.config("small")
    cs = sy.as_ctypes()
    L = ffi.lib()
    s = ffi.seeded_streams(S.seed_from_string(), 1)[0].copy()
    for _ in range(20):
        x, y, z = np.zeros(12), np.zeros(12), np.zeros(12)
        L.orc_rotqua(C.byref(cs), ffi.U(s), ffi.P(x), ffi.P(y), ffi.P(z))
        P0 = np.stack([sy.xq0, sy.yq0, sy.zq0], 1)
        P1 = np.stack([x, y, z], 1)
        assert np.max(np.abs(P1 @ P1.T - P0 @ P0.T)) < 1e-13


def test_yl0_addition_theorem():
    """sum_q w_q yl0(l, n.q) P_L'(m.q) = delta_LL' P_L(n.m) for L+L' <= 5."""
    L = ffi.lib()
    xq, yq, zq, wq = S.icosa12()
    Q = np.stack([xq, yq, zq], 1)
    rng = np.random.default_rng(3)
    leg = [lambda t: 1.0, lambda t: t, lambda t: 0.5 * (3 * t * t - 1), lambda t: 0.5 * (5 * t ** 3 - 3 * t)]
    for _ in range(5):
        n = rng.normal(size=3); n /= np.linalg.norm(n)
        m = rng.normal(size=3); m /= np.linalg.norm(m)
        for l in range(1, 4):
            for lp in range(0, 4):
                if (l - 1) + lp > 5:
                    continue
                s = sum(wq[q] * L.orc_yl0(l, float(n @ Q[q])) * leg[lp](float(m @ Q[q])) for q in range(12))
                ref = leg[l - 1](float(n @ m)) if lp == l - 1 else 0.0
                assert abs(s - ref) < 1e-10, (l, lp)


# ---------------------------------------------------- radial tables + AOs
def test_spline_tables_reproduce_contractions():
    sy = S.config("c22")
    cs = sy.as_ctypes()
    L = ffi.lib()
    r0, x = S.grid3()
    rng = np.random.default_rng(0)
    for it in range(sy.nctype):
        for irb, (al, co, l) in enumerate(sy.shells[it]):
            def R(r):
                return sum(S.gnorm(a, l) * c * np.exp(-a * r * r) for a, c in zip(al, co))
            def dR(r):
                return sum(-2 * a * r * S.gnorm(a, l) * c * np.exp(-a * r * r) for a, c in zip(al, co))
            wfv = np.zeros(4)
            for i in (3, 100, 900, 1500):  # nodes
                L.orc_splfit(C.byref(cs), it, irb, x[i], ffi.P(wfv), 1)
                assert abs(wfv[0] - sy.rwf[it, irb, i]) <= 1e-12 * max(1.0, abs(wfv[0]))
            for r in rng.uniform(0.01, min(6.0, sy.rmaxwf[it, irb]), size=30):
                L.orc_splfit(C.byref(cs), it, irb, r, ffi.P(wfv), 1)
                scale = max(abs(R(0.0)), 1.0)
                assert abs(wfv[0] - R(r)) < 2e-6 * scale
                assert abs(wfv[1] - dR(r)) < 2e-4 * scale * 10


def _ao_analytic(sy, pos):
    """Closed-form Cartesian Gaussian AOs: poly(x,y,z) * sum gnorm d exp(-a r^2)."""
    consts = {1: 0.28209479177387814, 2: 0.48860251190291992, 5: 0.630783130505040, 6: 1.092548430592079}
    mono = {1: (0, 0, 0), 2: (1, 0, 0), 3: (0, 1, 0), 4: (0, 0, 1), 5: (2, 0, 0), 6: (1, 1, 0), 7: (1, 0, 1),
            8: (0, 2, 0), 9: (0, 1, 1), 10: (0, 0, 2)}
    cst = {1: consts[1], 2: consts[2], 3: consts[2], 4: consts[2], 5: consts[5], 6: consts[6], 7: consts[6],
           8: consts[5], 9: consts[6], 10: consts[5]}
    out = []
    for ic in range(sy.ncent):
        it = sy.iwctype[ic]
        d = pos - sy.cent[ic]
        r2 = float(d @ d)
        for i in range(sy.nbastyp[it]):
            l = sy.iwlbas[it, i]
            al, co, ll = sy.shells[it][sy.iwrwf[it, i]]
            R = sum(S.gnorm(a, ll) * c * math.exp(-a * r2) for a, c in zip(al, co))
            px, py, pz = mono[l]
            out.append(cst[l] * d[0] ** px * d[1] ** py * d[2] ** pz * R)
    return np.array(out)


def test_basis_fns_vs_closed_form_and_fd():
    sy = S.config("c22")
    cs = sy.as_ctypes()
    L = ffi.lib()
    nb = sy.nbasis
    rng = np.random.default_rng(5)

    def phi_at(p, ider=2):
        rvec = (p[None, :] - sy.cent).copy()
        r = np.sqrt((rvec ** 2).sum(1))
        ph, dph, d2 = np.zeros(nb), np.zeros(3 * nb), np.zeros(nb)
        L.orc_basis_fns_point(C.byref(cs), ffi.P(rvec), ffi.P(r), ider, ffi.P(ph), ffi.P(dph), ffi.P(d2))
        return ph, dph.reshape(3, nb), d2

    for _ in range(6):
        p = sy.cent[rng.integers(sy.ncent)] + rng.normal(0, 0.8, 3)
        ph, dph, d2 = phi_at(p)
        ref = _ao_analytic(sy, p)
        assert np.max(np.abs(ph - ref)) < 1e-5 * max(1, np.max(np.abs(ref)))
        h = 1e-4
        g_fd = np.zeros((3, nb)); lap = np.zeros(nb)
        for k in range(3):
            e = np.zeros(3); e[k] = h
            fp, fm = phi_at(p + e)[0], phi_at(p - e)[0]
            g_fd[k] = (fp - fm) / (2 * h)
            lap += (fp - 2 * ph + fm) / h ** 2
        sc = max(1, np.max(np.abs(dph)))
        assert np.max(np.abs(g_fd - dph)) < 1e-6 * sc
        assert np.max(np.abs(lap - d2)) < 1e-3 * max(1, np.max(np.abs(d2)))


# ------------------------------------------------------------- Jastrow
@pytest.mark.parametrize("name", ["small", "c22"])
def test_jastrow_incremental_equals_full(name):
    sy = S.config(name)
    x = sy.initial_walkers(1, seed=2)[0]
    w = ffi.walker(sy)
    w.set_x(x)
    w.hpsi()
    rng = np.random.default_rng(1)
    for iel in range(sy.nelec):
        xn = x[iel] + rng.normal(0, 0.5, 3)
        _, psijn = w.psie(iel, xn)
        vjn = w.vjn()
        x2 = x.copy(); x2[iel] = xn
        w2 = ffi.walker(sy); w2.set_x(x2)
        fs, d2, fj = w2.jastrow_full()
        assert abs(psijn - fs) < 1e-12 * max(1, abs(fs))
        assert np.max(np.abs(vjn - fj)) < 1e-11 * max(1, np.max(np.abs(fj)))
        w.reject(iel)


def test_jastrow_gradient_laplacian_fd():
    sy = S.config("small_ae")
    x = sy.initial_walkers(1, seed=3)[0]
    w = ffi.walker(sy)

    def J(xx):
        w.set_x(xx)
        return w.jastrow_full()

    f0, d20, fj0 = J(x)
    h = 1e-4
    lap = 0.0
    for i in range(sy.nelec):
        for k in range(3):
            xp = x.copy(); xp[i, k] += h
            xm = x.copy(); xm[i, k] -= h
            fp, fm = J(xp)[0], J(xm)[0]
            assert abs((fp - fm) / (2 * h) - fj0[i, k]) < 1e-6 * max(1, abs(fj0[i, k]))
            lap += (fp - 2 * f0 + fm) / h ** 2
    assert abs(lap - d20) < 1e-4 * max(1, abs(d20))


@pytest.mark.parametrize("pair,cusp", [((0, 1), 0.25), ((0, 5), 0.5)])
def test_jastrow_ee_cusp(pair, cusp):
    """dJ/dr_ij -> sspinn * b1 at coalescence (e-e-n terms made cusp-free)."""
    sy = S.config("small_ae")  # nup = 5: (0,1) parallel, (0,5) antiparallel
    x = sy.initial_walkers(1, seed=4)[0]
    w = ffi.walker(sy)
    i, j = pair
    n = np.array([0.3, -0.5, 0.8]); n /= np.linalg.norm(n)
    # J = cusp * |r_ij| + smooth: the symmetric second difference isolates the kink
    eps = 1e-5
    vals = []
    for s in (1.0, -1.0, 0.0):
        xx = x.copy(); xx[j] = xx[i] + s * eps * n
        w.set_x(xx)
        vals.append(w.jastrow_full()[0])
    slope = (vals[0] + vals[1] - 2 * vals[2]) / (2 * eps)
    assert abs(slope - cusp) < 1e-4


def test_jastrow_en_cusp_and_asymptote():
    sy = S.config("small_ae")
    x = sy.initial_walkers(1, seed=6)[0]
    w = ffi.walker(sy)
    # e-n cusp: a1 = -Z is carried by the Jastrow; C terms must not add to it
    ic = 0
    n = np.array([0.2, 0.9, -0.4]); n /= np.linalg.norm(n)
    eps = 1e-5
    vals = []
    for s in (1.0, -1.0, 0.0):
        xx = x.copy(); xx[0] = sy.cent[ic] + s * eps * n
        w.set_x(xx)
        vals.append(w.jastrow_full()[0])
    slope = (vals[0] + vals[1] - 2 * vals[2]) / (2 * eps)
    assert abs(slope - sy.a4[sy.iwctype[ic], 0]) < 1e-4
    # all particles far apart: every term -> its subtracted asymptote -> J -> 0
    xx = np.array([[1e4 * (k + 1), -1e4 * k, 3e3 * k] for k in range(sy.nelec)], dtype=float)
    w.set_x(xx)
    assert abs(w.jastrow_full()[0]) < 1e-12


# ------------------------------------------------------- Slater / SM
def test_sherman_morrison_equals_fresh_lu():
    sy = S.config("c22")
    x = sy.initial_walkers(1, seed=8)[0]
    w = ffi.walker(sy); w.set_x(x); h0 = w.hpsi()
    rng = np.random.default_rng(2)
    for iel in (0, 5, 11, 21):
        xn = x[iel] + rng.normal(0, 0.3, 3)
        pdn, pjn = w.psie(iel, xn)
        w.accept(iel)
        x[iel] = xn
        w2 = ffi.walker(sy); w2.set_x(x); h2 = w2.hpsi()
        assert abs(pdn / h2["psid"] - 1) < 1e-12
        assert np.max(np.abs(w.slmi() - w2.slmi())) < 1e-10 * np.max(np.abs(w2.slmi()))


def test_multidet_moves_equal_fresh_evaluation():
    """Per-move multideterminant psi and drift (every determinant rebuilt with
    the moved row) against a fresh hpsi at the moved configuration."""
    sy = S.config("c22_md")
    x = sy.initial_walkers(1, seed=8)[0]
    w = ffi.walker(sy); w.set_x(x); w.hpsi()
    rng = np.random.default_rng(3)
    for iel in (0, 5, 11, 21):
        xn = x[iel] + rng.normal(0, 0.3, 3)
        pdn, pjn = w.psie(iel, xn)
        vn = w.grad(iel, 0)
        w.accept(iel)
        x[iel] = xn
        w2 = ffi.walker(sy); w2.set_x(x); h2 = w2.hpsi()
        assert abs(pdn / h2["psid"] - 1) < 1e-11
        assert np.max(np.abs(vn - w2.grad(iel, 1))) < 1e-10 * max(1.0, np.max(np.abs(vn)))
        assert np.max(np.abs(w.grad(iel, 1) - vn)) < 1e-12 * max(1.0, np.max(np.abs(vn)))


def test_multidet_phase_of_reordered_reference():
    """Two determinants with the reference's orbitals, the second with two
    slots of each spin exchanged: psi_D = (c0 + c1) D_ref (two exchanges) and
    every derivative of ln psi_D equals the single determinant's."""
    sy1 = S.config("small")
    sy = S.config("small")
    iw = np.repeat(sy.iworbd, 2, axis=0)
    iw[1, [0, 1]] = iw[1, [1, 0]]
    iw[1, [5, 7]] = iw[1, [7, 5]]
    sy.iworbd = iw
    sy.cdet = np.array([0.7, 0.2])
    x = sy.initial_walkers(1, seed=4)[0]
    a = ffi.walker(sy1); a.set_x(x); ha = a.hpsi()
    b = ffi.walker(sy); b.set_x(x); hb = b.hpsi()
    assert abs(hb["psid"] / (0.9 * ha["psid"]) - 1) < 1e-12
    assert abs(hb["energy"] - ha["energy"]) < 1e-10 * abs(ha["energy"])
    assert np.max(np.abs(b.ddx()[0] - a.ddx()[0])) < 1e-10 * np.max(np.abs(a.ddx()[0]))
    iw[1, [0, 1]] = iw[1, [1, 0]]   # one exchange left: odd permutation
    sy.iworbd = iw
    c = ffi.walker(sy); c.set_x(x); hc = c.hpsi()
    assert abs(hc["psid"] / (0.5 * ha["psid"]) - 1) < 1e-12


# ------------------------------------------------------ drift and E_L
def _lnpsi(w, xx):
    w.set_x(xx)
    h = w.hpsi()
    return math.log(abs(h["psid"])) + h["psij"], h


@pytest.mark.parametrize("name", ["small_ae", "small", "small_ae_md", "small_md"])
def test_drift_and_kinetic_vs_finite_differences(name):
    sy = S.config(name)
    x = sy.initial_walkers(1, seed=12)[0]
    w = ffi.walker(sy)
    f0, h0 = _lnpsi(w, x)
    drift = np.array([w.grad(i, 1) for i in range(sy.nelec)])
    hh = 1e-4
    lap = 0.0
    g2 = 0.0
    for i in range(sy.nelec):
        for k in range(3):
            xp = x.copy(); xp[i, k] += hh
            xm = x.copy(); xm[i, k] -= hh
            fp, fm = _lnpsi(w, xp)[0], _lnpsi(w, xm)[0]
            g = (fp - fm) / (2 * hh)
            assert abs(g - drift[i, k]) < 1e-6 * max(1, abs(g))
            lap += (fp - 2 * f0 + fm) / hh ** 2
            g2 += g * g
    ekin_fd = -0.5 * (lap + g2)
    assert abs(ekin_fd - h0["ekin"]) < 1e-4 * max(1, abs(h0["ekin"]))


@pytest.mark.parametrize("name", ["small", "small_md"])
def test_nonlocal_energy_by_brute_force(name):
    """vpsp_det re-evaluated from full psi values at the quadrature points."""
    sy = S.config(name)
    x = sy.initial_walkers(1, seed=21)[0]
    w = ffi.walker(sy)
    f0, h0 = _lnpsi(w, x)
    parts = w.parts()
    psi0 = h0["psid"] * math.exp(h0["psij"])
    L = ffi.lib()
    cs = sy.as_ctypes()
    xq, yq, zq, wq = S.icosa12()
    tot = [0.0, 0.0]
    for i in range(sy.nelec):
        for ic in range(sy.ncent):
            it = int(sy.iwctype[ic]); lp = int(sy.lpot[it])
            rv = x[i] - sy.cent[ic]; r = float(np.linalg.norm(rv))
            vps = []
            for l in range(1, lp):
                v, dv = np.zeros(1), np.zeros(1)
                L.orc_gauss_pot(C.byref(cs), max(r, 1e-10), l, it, ffi.P(v), ffi.P(dv))
                vps.append(v[0])
            if not any(abs(v) > 1e-4 for v in vps):
                continue
            for q in range(12):
                qv = np.array([xq[q], yq[q], zq[q]])
                costh = float(rv @ qv) / r
                xx = x.copy(); xx[i] = sy.cent[ic] + r * qv
                _, hq = _lnpsi(w, xx)
                ratio = hq["psid"] * math.exp(hq["psij"]) / psi0
                term = sum(L.orc_yl0(l, costh) * vps[l - 1] for l in range(1, lp))
                tot[0 if i < sy.nup else 1] += term * wq[q] * ratio
    assert abs(tot[0] - parts["vpsp_up"]) < 1e-10 * max(1, abs(tot[0]))
    assert abs(tot[1] - parts["vpsp_dn"]) < 1e-10 * max(1, abs(tot[1]))


def test_pe_local_all_electron():
    sy = S.config("small_ae")
    x = sy.initial_walkers(1, seed=1)[0]
    w = ffi.walker(sy); w.set_x(x); w.hpsi()
    pe = 0.0
    for i in range(1, sy.nelec):
        for j in range(i):
            pe += 1 / np.linalg.norm(x[i] - x[j])
    for i in range(sy.nelec):
        for ic in range(sy.ncent):
            pe -= sy.znuc[sy.iwctype[ic]] / np.linalg.norm(x[i] - sy.cent[ic])
    for i in range(1, sy.ncent):
        for j in range(i):
            pe += sy.znuc[sy.iwctype[i]] * sy.znuc[sy.iwctype[j]] / np.linalg.norm(sy.cent[i] - sy.cent[j])
    assert abs(pe - w.parts()["pe_local"]) < 1e-12 * abs(pe)


# ------------------------------------------------ pins added for gauss_pot,
# the ECP local channel, rotqua and the e-e-n coefficient order
def _bfd_system():
    """A system carrying the BFD carbon and hydrogen pseudopotentials read from
    the reference's CI files (tests/golden/champ_inputs/C4H6_ci44/pool)."""
    from champ_amd import champ_io as io
    sy = S.config("small")
    pool = os.path.join(HERE, "golden", "champ_inputs", "C4H6_ci44", "pool")
    for it, atom in enumerate(("C", "H")):
        lp, comps = io.read_gauss_ecp(os.path.join(pool, f"BFD.gauss_ecp.dat.{atom}"))
        sy.lpot[it] = lp
        sy.necp_term[it] = 0
        for idx, terms in comps.items():
            sy.necp_term[it, idx] = len(terms)
            for g, (cf, pw, ex) in enumerate(terms):
                sy.ecp_coef[it, idx, g], sy.necp_power[it, idx, g], sy.ecp_exponent[it, idx, g] = cf, pw, ex
    sy.znuc[:] = [4.0, 1.0]
    return sy


def _vps(L, cs, sy, r, l, it):
    v, dv = np.zeros(1), np.zeros(1)
    L.orc_gauss_pot(C.byref(cs), float(r), l, it, ffi.P(v), ffi.P(dv))
    return v[0], dv[0]


def test_gauss_pot_bfd_closed_forms():
    """BFD pseudopotentials are finite at the nucleus by construction: the
    local channel's c r^-1 e^(-a r^2) term (power 1) cancels -Z/r, so
    V_loc(0) = the coefficient of its power-2 term (C: -19.17537323,
    H: -8.228005709676), and a non-local channel of one power-2 term equals
    its coefficient at r = 0 (C s: 22.55164191).  Beyond a r^2 > 700 every
    Gaussian is dropped and V_loc = -Z/r exactly.  dV/dr against central
    differences.  A wrong power offset (n - 2), a missing -Z/r or a local
    component stored at the wrong index fails one of these."""
    sy = _bfd_system()
    cs = sy.as_ctypes()
    L = ffi.lib()
    for it, z, v0 in ((0, 4.0, -19.17537323), (1, 1.0, -8.228005709676)):
        lp = int(sy.lpot[it])
        r = 1e-7
        v, _ = _vps(L, cs, sy, r, lp, it)
        assert abs((v - z / r) - v0) < 1e-5, (it, v - z / r)
        r = 14.0  # smallest exponent 3.94: 3.94 * 196 > 700
        v, dv = _vps(L, cs, sy, r, lp, it)
        assert v == 0.0 and dv == 0.0  # gauss_pot alone; getvps adds -Z/r
        for r in (0.3, 0.9, 2.2):
            h = 1e-5
            vp, _ = _vps(L, cs, sy, r + h, lp, it)
            vm, _ = _vps(L, cs, sy, r - h, lp, it)
            _, dv = _vps(L, cs, sy, r, lp, it)
            assert abs((vp - vm) / (2 * h) - dv) < 1e-7 * max(1.0, abs(dv))
    v, _ = _vps(L, cs, sy, 1e-9, 1, 0)
    assert abs(v - 22.55164191) < 1e-6


def test_pe_local_ecp_closed_form():
    """pe_local on a pseudopotential system: nuclear repulsion, electron
    repulsion and, per (electron, centre), the local channel
    sum_k c_k r^(n_k - 2) exp(-a_k r^2) - Z/r, evaluated here from the BFD
    files' own numbers (nonloc_pot.f90:44-63, readps_gauss.f90:6-92)."""
    from champ_amd import champ_io as io
    sy = _bfd_system()
    x = sy.initial_walkers(1, seed=2)[0]
    w = ffi.walker(sy)
    w.set_x(x)
    w.set_quad(sy.xq0, sy.yq0, sy.zq0)
    w.hpsi()
    pool = os.path.join(HERE, "golden", "champ_inputs", "C4H6_ci44", "pool")
    local = {}
    for it, atom in enumerate(("C", "H")):
        lp, comps = io.read_gauss_ecp(os.path.join(pool, f"BFD.gauss_ecp.dat.{atom}"))
        local[it] = comps[lp - 1]
    pe = 0.0
    for i in range(1, sy.nelec):
        for j in range(i):
            pe += 1 / np.linalg.norm(x[i] - x[j])
    for i in range(1, sy.ncent):
        for j in range(i):
            pe += sy.znuc[sy.iwctype[i]] * sy.znuc[sy.iwctype[j]] / np.linalg.norm(sy.cent[i] - sy.cent[j])
    for i in range(sy.nelec):
        for ic in range(sy.ncent):
            it = int(sy.iwctype[ic])
            r = float(np.linalg.norm(x[i] - sy.cent[ic]))
            pe += sum(cf * r ** (pw - 2) * math.exp(-ex * r * r) for cf, pw, ex in local[it] if ex * r * r <= 700)
            pe -= sy.znuc[it] / r
    assert abs(pe - w.parts()["pe_local"]) < 1e-12 * abs(pe)


def test_rotqua_known_answer():
    """rotqua from the walker's stream: the first accepted pair (x1, x2) of
    uniforms in the unit disk gives Marsaglia's point X on the sphere, the
    second accepted pair (u1, u2) an angle in the plane orthogonal to X:
    Y = u e_theta + v e_phi, Z = v e_theta - u e_phi with e_theta, e_phi the
    spherical unit vectors at X (rotqua.f90:363-419).  The rule's points q map
    to q_x X + q_y Y + q_z Z.  (Y, Z, X) is left-handed: det = -1, a
    reflection that the rule's inversion symmetry leaves harmless.  Checked
    on the 12-point rule for 20 draws."""
    sy = S.config("small")
    cs = sy.as_ctypes()
    L = ffi.lib()
    st = ffi.seeded_streams(S.seed_from_string(), 1)[0].copy()
    for _ in range(20):
        s_py = st.copy()
        u = lambda: L.orc_random_dp(ffi.U(s_py))  # noqa: E731
        while True:
            a, b = 1 - 2 * u(), 1 - 2 * u()
            if a * a + b * b < 1:
                break
        s2 = a * a + b * b
        X = np.array([2 * a * math.sqrt(1 - s2), 2 * b * math.sqrt(1 - s2), 1 - 2 * s2])
        th, ph = math.acos(X[2]), math.atan2(X[1], X[0])
        e_th = np.array([math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph), -math.sin(th)])
        e_ph = np.array([-math.sin(ph), math.cos(ph), 0.0])
        while True:
            p, q = 2 * u() - 1, 2 * u() - 1
            if p * p + q * q < 1:
                break
        n = math.hypot(p, q)
        p, q = p / n, q / n
        Y = p * e_th + q * e_ph
        Z = q * e_th - p * e_ph
        xq, yq, zq = np.zeros(12), np.zeros(12), np.zeros(12)
        L.orc_rotqua(C.byref(cs), ffi.U(st), ffi.P(xq), ffi.P(yq), ffi.P(zq))
        expect = np.outer(sy.xq0, X) + np.outer(sy.yq0, Y) + np.outer(sy.zq0, Z)
        assert np.max(np.abs(np.stack([xq, yq, zq], 1) - expect)) < 1e-12
        assert np.array_equal(st, s_py)  # the same number of draws
        assert abs(np.linalg.det(np.stack([X, Y, Z])) + 1) < 1e-12


# (k, l, m) of each c parameter for nordc = 4, from CHAMP's published form of
# f_een (docs/users/preparation/jastrow.md:60-66: p = 2..nordc, k = p-1..0,
# l = lmax..0, m = (p-k-l)/2 when integer; lmax = p-k, or p-k-2 when k = 0
# so that no pure e-n term u^0 (x^p + y^p) appears)
EEN_NORDC4 = [(1, 1, 0), (0, 0, 1),
              (2, 1, 0), (1, 2, 0), (1, 0, 1), (0, 1, 1),
              (3, 1, 0), (2, 2, 0), (2, 0, 1), (1, 3, 0), (1, 1, 1), (0, 2, 1), (0, 0, 2)]


def test_een_coefficient_order_one_hot():
    """Two electrons, one centre, a = b = 0 and a one-hot c: the oracle's log
    Jastrow is exactly c u^k (x^l + y^l) (x y)^m with u, x, y the switched
    distances exp(-kappa r) of the pair and of each electron to the centre."""
    base = S.make_system("een2", 1, 0, 2, "sp", "s", nloc=0, nordc=4, seed=5)
    assert S.een_terms(4) == EEN_NORDC4 and S.nterms4(4) == 13
    x = np.array([[0.3, -0.2, 0.5], [-0.4, 0.6, 0.1]])
    kap = base.scalek
    u = math.exp(-kap * np.linalg.norm(x[0] - x[1]))
    xa = math.exp(-kap * np.linalg.norm(x[0] - base.cent[0]))
    ya = math.exp(-kap * np.linalg.norm(x[1] - base.cent[0]))
    for t, (k, l, m) in enumerate(EEN_NORDC4):
        sy = S.System(**{**base.__dict__, "a4": np.zeros_like(base.a4), "b": np.zeros_like(base.b),
                         "c": np.zeros_like(base.c)})
        sy.c[0, t] = 0.7
        w = ffi.walker(sy)
        w.set_x(x)
        psij = w.hpsi()["psij"]
        assert abs(psij - 0.7 * u ** k * (xa ** l + ya ** l) * (xa * ya) ** m) < 1e-14, (t, (k, l, m))


# ----------------------------------------------------------------- metrop6
@pytest.mark.parametrize("a", [1.5, 2.5])
def test_gammai_matches_scipy(a):
    """gammai (gammai.f90): the lower incomplete gamma function for x < a+4
    (iflag 1) and minus the upper one beyond (iflag -1), with its fixed term
    counts, against scipy's regularised functions times Gamma(a)."""
    from scipy import special
    L = ffi.lib()
    fl = C.c_int()
    G = special.gamma(a)
    for x in np.concatenate([np.linspace(1e-3, 40.0, 401), [a + 4 - 1e-9, a + 4, 60.0, 100.0]]):
        g = L.orc_gammai(a, float(x), float(x ** a * np.exp(-x)), C.byref(fl))
        ref = special.gammainc(a, x) * G if x < a + 4 else -special.gammaincc(a, x) * G
        assert fl.value == (1 if x < a + 4 else -1)
        assert abs(g - ref) <= 1e-12 * abs(ref), (a, x, g, ref)


def test_metrop6_radial_area_matches_quadrature():
    """The normalisation metrop6 forms from incomplete gamma functions
    (metrop_mov1_slat.f90:256-293) equals the integral of
    sqrt(r)|1 + co r| exp(-zeta r) over [rmin/5, 5 rmin] by adaptive
    quadrature, divided by the density at rmin; cases with the zero of
    1 + co r inside and outside the interval."""
    from scipy import integrate
    L = ffi.lib()
    rng = np.random.default_rng(1)
    inside = 0
    for _ in range(300):
        zeta, co, rmin = rng.uniform(0.05, 6.0), rng.uniform(-3.0, 3.0), rng.uniform(0.05, 4.0)
        rb, rt = rmin / 5, rmin * 5
        bot = math.sqrt(rmin) * abs(1 + co * rmin) * math.exp(-zeta * rmin)
        if bot == 0.0:
            continue
        pts = [-1 / co] if rb < -1 / co < rt else None
        inside += pts is not None
        v, _ = integrate.quad(lambda r: math.sqrt(r) * abs(1 + co * r) * math.exp(-zeta * r), rb, rt, points=pts,
                              epsabs=0, epsrel=1e-13, limit=200)
        a = L.orc_m6_area(zeta, co, rb, rt, bot)
        assert abs(a - v / bot) <= 1e-12 * (v / bot), (zeta, co, rmin)
    assert inside > 20
