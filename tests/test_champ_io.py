"""CHAMP input readers on two of the reference's own CI cases (CPU only), and
the oracle run on them against the energies those CI scripts assert.

Fixtures: tests/golden/champ_inputs (made by make_fixtures.py from
tests/CI_test/VMC-H2 and tests/CI_test/VMC-C4H6-ci44_pVQZ_20_dets_12_csf_noopt).
The reference's comparison (tools/compare_value.py:14-28) accepts a value
within 2 sqrt(err^2 + err_ref^2) of the reference energy; the same rule is
applied here to the oracle's VMC energy over the same trial functions.
"""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from champ_amd import champ_io as io
from champ_amd import system as S
from oracle import ffi

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden", "champ_inputs")


def load_case(name: str):
    """The CI case's System (the butadiene orbitals converted from CHAMP's
    older AO order first, as the case's conversion script does)."""
    if name == "H2":
        return io.load_input(os.path.join(G, "H2", "revised_vmc.inp"), name="H2")
    d = os.path.join(G, "C4H6_ci44")
    types, _, iwc, _ = io.read_molecule(os.path.join(d, "pool", "champ_v3_butadiene_champ_v2_gamess_ci44_pVQZ_1_geom.xyz"), 4)
    coef = io.convert_lcao_order(io.read_lcao(os.path.join(d, "conversion", "champ_v2_gamess_ci44_pVQZ_cart.lcao")),
                                 os.path.join(d, "conversion", "champ_v2_gamess_ci44_pVQZ.bfinfo"),
                                 os.path.join(d, "pool", "champ_v3_champ_v2_gamess_ci44_pVQZ_basis_pointers.bfinfo"),
                                 [types[i] for i in iwc], types)
    return io.load_input(os.path.join(d, "vmc_noopt_ci44.inp"), orbitals=coef, name="C4H6_ci44")


def reference_energy(name: str, cores: int = 1):
    with open(os.path.join(G, name, "reference_energies.json")) as fh:
        for e in json.load(fh)["total_E"]:
            if e["cores"] == cores:
                return e["energy"], e["error"]
    raise KeyError(cores)


@pytest.fixture(scope="module")
def h2():
    return load_case("H2")


@pytest.fixture(scope="module")
def c4h6():
    return load_case("C4H6_ci44")


def test_h2_tables(h2):
    assert (h2.nelec, h2.nup, h2.ncent, h2.nctype, h2.nbasis, h2.norb, h2.nloc) == (2, 1, 2, 1, 12, 1, 0)
    assert h2.znuc[0] == 1.0 and (h2.norda, h2.nordb, h2.nordc) == (5, 5, 0) and h2.scalek == 0.6
    assert list(h2.b[0]) == [0.5, 0.5, 0, 0, 0, 0] and not np.any(h2.a4)
    # s s s px py pz on radial functions 1 2 3 4 4 4 (champ_v3_sto_cvb1_basis_pointers.bfinfo)
    assert list(h2.iwlbas[0]) == [1, 1, 1, 2, 3, 4] and list(h2.iwrwf[0]) == [0, 1, 2, 3, 3, 3]
    # grid type 3 as read_bas_num.f90:146 forms it: r0 = 20 / (1.003^1999 - 1)
    assert h2.nr == 2000 and abs(h2.grid_r0 - 20.0 / (1.003 ** 1999 - 1)) < 1e-15


def test_c4h6_tables(c4h6):
    sy = c4h6
    assert (sy.nelec, sy.nup, sy.ncent, sy.nctype, sy.nbasis, sy.norb, sy.nloc) == (22, 11, 10, 2, 426, 13, 4)
    assert list(sy.znuc) == [4.0, 1.0]  # C and H with the BFD cores removed
    assert len(sy.wq) == 6  # CHAMP's default nquad (parser.f90:578)
    # carbon: local component (3 terms) stored at lpot = 2, s non-local (1 term) at 1
    assert list(sy.lpot) == [2, 1] and sy.necp_term[0, 1] == 3 and sy.necp_term[0, 0] == 1 and sy.necp_term[1, 0] == 3
    assert sy.necp_power[0, 1, 0] == 1 and abs(sy.ecp_exponent[0, 1, 0] - 8.35973821) < 1e-15


def test_c4h6_determinant_coefficients_from_csfs(c4h6):
    """cdet rebuilt from the CSF coefficients and map equals the determinant
    coefficients the file also lists (written to 8 decimals)."""
    path = os.path.join(G, "C4H6_ci44", "champ_v2_gamess_ci44_pVQZ_cart.det")
    with open(path) as fh:
        lines = fh.read().splitlines()
    k = next(i for i, l in enumerate(lines) if l.split()[:1] == ["determinants"])
    listed = np.array([float(v) for v in lines[k + 1].split()])
    assert np.max(np.abs(c4h6.cdet - listed)) < 1e-6
    # orbitals 0-based, determinant 0 the reference: 11 up and 11 down orbitals 0..10
    assert list(c4h6.iworbd[0]) == list(range(11)) * 2


def test_lcao_order_conversion_is_a_permutation(c4h6):
    d = os.path.join(G, "C4H6_ci44")
    old = io.read_lcao(os.path.join(d, "conversion", "champ_v2_gamess_ci44_pVQZ_cart.lcao"))
    # each orbital's coefficients are moved between AO slots, never scaled
    for o in range(13):
        assert np.array_equal(np.sort(old[o]), np.sort(c4h6.coef[o]))
    # an s function is its own component: the first four AOs of every atom stay put
    assert np.array_equal(old[:13, :4], c4h6.coef[:, :4])


def test_numerical_radial_tables(h2):
    """The spline through the file's nodes reproduces them (splfit at a node),
    and the tail fit continues the last nodes."""
    L = ffi.lib()
    cs = h2.as_ctypes()
    r0, x = S.grid3(h2.nr, h2.grid_arg, 20.0)
    for irb in range(h2.nrbas[0]):
        for i in (5, 400, 1500, 1990):
            wfv = np.zeros(4)
            L.orc_splfit(__import__("ctypes").byref(cs), 0, irb, float(x[i]), ffi.P(wfv), 0)
            assert abs(wfv[0] - h2.rwf[0, irb, i]) <= 1e-10 * max(1.0, abs(h2.rwf[0, irb, i])), (irb, i)


def _oracle_vmc(sy, nw, neq, nsw, tau, imetro=1):
    cs = sy.as_ctypes()
    x = sy.initial_walkers(nw, seed=1)
    st = ffi.seeded_streams(S.seed_from_string(), nw)

    def run(w):
        wk = ffi.walker(sy, cs)
        wk.set_x(x[w])
        wk.set_rng(st[w])
        wk.hpsi()
        sweep = (lambda: wk.sweep(tau)) if imetro == 1 else wk.sweep6
        for _ in range(neq):
            sweep()
        es = np.zeros(nsw)
        for k in range(nsw):
            sweep()
            es[k] = wk.scalars()["eold"]
        return es

    with ThreadPoolExecutor(8) as ex:
        E = np.array(list(ex.map(run, range(nw))))
    # blocks of 10 sweeps of the whole population (acuest blocking)
    blk = E.reshape(nw, nsw // 10, 10).mean(axis=(0, 2))
    return float(E.mean()), float(blk.std() / np.sqrt(len(blk)))


@pytest.mark.parametrize("name,nw,neq,nsw,tau", [("H2", 64, 100, 500, 0.5), ("C4H6_ci44", 32, 40, 100, 0.2)])
def test_oracle_vmc_energy_matches_reference_ci(name, nw, neq, nsw, tau, h2, c4h6):
    """The oracle's VMC energy over the CI case's trial function agrees with
    the energy the reference's CI asserts for it (metrop1 and CHAMP's default
    metrop6 sample the same |psi|^2)."""
    sy = h2 if name == "H2" else c4h6
    e, err = _oracle_vmc(sy, nw, neq, nsw, tau)
    eref, eref_err = reference_energy(name)
    print(f"[{name}] oracle {e:.5f} +- {err:.5f}; reference {eref} +- {eref_err}")
    assert abs(e - eref) <= 2.0 * np.hypot(err, eref_err)


@pytest.mark.parametrize("name,nw,neq,nsw", [("H2", 64, 100, 500), ("C4H6_ci44", 32, 40, 100)])
def test_oracle_metrop6_energy_matches_reference_ci(name, nw, neq, nsw, h2, c4h6):
    """The same with the oracle's metrop6 (metrop_mov1_slat.f90, CHAMP's
    default mover imetro=6 at its default deltar 5, deltat 1, fbias 1): the
    mover these CI energies were produced with."""
    sy = h2 if name == "H2" else c4h6
    e, err = _oracle_vmc(sy, nw, neq, nsw, 0.0, imetro=6)
    eref, eref_err = reference_energy(name)
    print(f"[{name} metrop6] oracle {e:.5f} +- {err:.5f}; reference {eref} +- {eref_err}")
    assert abs(e - eref) <= 2.0 * np.hypot(err, eref_err)
