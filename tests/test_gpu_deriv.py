"""Parameter derivatives d ln psi / d p on the GPU (champ_gpu_deriv_psi: the
gvalue of deriv_jastrow4.f90 for the linear Jastrow parameters and the
determinant ratios of optci.f90) against central finite differences of the
oracle's ln |psi| with the parameter moved in the system tables, and the
stochastic-reconfiguration sums (champ_gpu_sr_sums) against the same sums of
those derivatives."""
import numpy as np
import pytest

from champ_amd import system as S
from champ_amd.engine import DERIV_CI, DERIV_JAS_EE, DERIV_JAS_EEN, DERIV_JAS_EN, Engine
from oracle import ffi
from tests.gpu_helpers import ENGINE_LIB

pytestmark = pytest.mark.gpu

ALL = DERIV_JAS_EN | DERIV_JAS_EE | DERIV_JAS_EEN | DERIV_CI


def _params(sy):
    """(array attribute, index) of each parameter in the engine's order"""
    out = []
    for it in range(sy.nctype):
        for o in range(2, sy.norda + 1):
            out.append(("a4", (it, o)))
    for isb in range(sy.nspin2b):
        for o in range(2, sy.nordb + 1):
            out.append(("b", (isb, o)))
    for it in range(sy.nctype):
        for t in range(sy.c.shape[1]):
            out.append(("c", (it, t)))
    if sy.iworbd.shape[0] > 1:
        for k in range(sy.iworbd.shape[0]):
            out.append(("cdet", (k,)))
    return out


def _lnpsi(sy, x):
    cs = sy.as_ctypes()
    out = []
    for xw in x:
        w = ffi.walker(sy, cs)
        w.set_x(xw)
        h = w.hpsi()
        out.append(np.log(abs(h["psid"])) + h["psij"])
    return np.array(out)


@pytest.mark.parametrize("name", ["small", "small_md"])
def test_deriv_psi_matches_finite_differences(require_gpu, name):
    sy = S.config(name)
    nw = 3
    x = sy.initial_walkers(nw, seed=17)
    eng = Engine(sy, nw, lib_path=ENGINE_LIB)
    eng.set_walkers(x)
    eng.hpsi()
    o = eng.deriv_psi(ALL)
    params = _params(sy)
    assert o.shape == (nw, len(params))
    for k, (attr, idx) in enumerate(params):
        arr = getattr(sy, attr)
        p0 = float(arr[idx])
        h = 1e-5 * max(1.0, abs(p0))
        arr[idx] = p0 + h
        lp = _lnpsi(sy, x)
        arr[idx] = p0 - h
        lm = _lnpsi(sy, x)
        arr[idx] = p0
        fd = (lp - lm) / (2 * h)
        scale = max(1.0, float(np.max(np.abs(fd))))
        err = float(np.max(np.abs(o[:, k] - fd))) / scale
        assert err < 1e-6, (name, attr, idx, o[:, k], fd)
    print(f"[deriv {name}] {len(params)} parameters match finite differences", flush=True)


@pytest.mark.parametrize("name", ["c22_md"])
def test_sr_sums_are_the_derivative_sums(require_gpu, name):
    sy = S.config(name)
    nw = 64
    eng = Engine(sy, nw, lib_path=ENGINE_LIB)
    eng.set_walkers(sy.initial_walkers(nw, seed=4))
    eng.seed_rng(S.seed_from_string())
    eng.hpsi()
    eng.metrop1(2, 0.3)
    o = eng.deriv_psi(ALL)
    s, os_, eo, es = eng.sr_sums(ALL)
    e = eng.scalars()["eold"]
    assert np.allclose(s, o.T @ o, rtol=1e-12, atol=1e-12 * np.max(np.abs(o.T @ o)))
    assert np.allclose(os_, o.sum(axis=0), rtol=1e-12, atol=1e-13)
    assert np.allclose(eo, e @ o, rtol=1e-12, atol=1e-12 * np.max(np.abs(e @ o)))
    assert abs(es - e.sum()) <= 1e-12 * abs(e.sum())
