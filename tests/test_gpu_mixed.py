"""The mixed-precision path (CHAMP_PREC_MIXED) against the FP64 path on
identical walker configurations: per-walker local energy and log psi
differences, and a short metrop1 run with the FP64 refresh of the Slater
inverse every 3 moves (set_lu_every; the default 80 only acts above 80
electrons, which tests/test_gpu_configs.py covers at 160 electrons).

The bound is relative and loose: FP32 carries about 7 digits, and the
synthetic systems' local energies (random wavefunctions) reach hundreds of
hartree, so the absolute E_L error is not below 1e-6 Ha here.  The measured
errors are printed; bench.py reports them on the full-size system.
"""
import numpy as np
import pytest

from champ_amd import system as S
from champ_amd.engine import Engine, PREC_FP64, PREC_MIXED
from tests.gpu_helpers import ENGINE_LIB

pytestmark = pytest.mark.gpu


def _pair(name, nwalk, seed):
    sy = S.config(name)
    x0 = sy.initial_walkers(nwalk, seed=seed)
    out = []
    for prec in (PREC_FP64, PREC_MIXED):
        e = Engine(sy, nwalk, precision=prec, lib_path=ENGINE_LIB)
        e.set_walkers(x0)
        e.seed_rng(S.seed_from_string())
        out.append(e)
    return sy, out


@pytest.mark.parametrize("name,nwalk", [("small", 8), ("c22", 8), ("small_md", 6), ("c22_md", 4)])
def test_mixed_hpsi_close_to_fp64(require_gpu, name, nwalk):
    sy, (e64, e32) = _pair(name, nwalk, seed=11)
    g64, g32 = e64.hpsi(), e32.hpsi()
    de = np.abs(g32["energy"] - g64["energy"])
    rel_e = de / np.maximum(1.0, np.abs(g64["energy"]))
    lnpsi64 = np.log(np.abs(g64["psid"])) + g64["psij"]
    lnpsi32 = np.log(np.abs(g32["psid"])) + g32["psij"]
    dlp = np.abs(lnpsi32 - lnpsi64)
    print(f"[mixed {name}] |dE_L| max {de.max():.3e} Ha, rel {rel_e.max():.3e}; |dln psi| max {dlp.max():.3e}", flush=True)
    assert np.all(np.isfinite(g32["energy"]))
    assert rel_e.max() < 1e-4
    assert dlp.max() < 1e-4
    assert not np.any(e32.status())


@pytest.mark.parametrize("name", ["small", "small_md"])
def test_mixed_metrop1_runs(require_gpu, name):
    sy, (e64, e32) = _pair(name, 16, seed=13)
    e64.hpsi(); e32.hpsi()
    e32.set_lu_every(3)
    s64 = e64.metrop1(3, tau=0.3)
    s32 = e32.metrop1(3, tau=0.3)
    print(f"[mixed metrop1 {name}] acc fp64 {s64['accepted'] / s64['proposed']:.3f} mixed {s32['accepted'] / s32['proposed']:.3f}",
          flush=True)
    assert s32["nflagged"] == 0
    assert np.isfinite(s32["esum"])
    # same seeds and near-identical ratios: the acceptance counts agree closely
    assert abs(s32["accepted"] - s64["accepted"]) <= 0.05 * s64["accepted"]
