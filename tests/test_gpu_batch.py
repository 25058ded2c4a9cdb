"""champ_gpu_metrop1_batch: several populations (different molecules, walker
counts and precisions) swept at once, each on its own stream.  Every
population must evolve exactly as under its own champ_gpu_metrop1: walkers,
per-walker energies and acceptance counts bit-identical; the estimator sums
are accumulated across walkers by atomic adds, whose order is not fixed, so
they agree to rounding."""
import numpy as np
import pytest

from champ_amd.engine import PREC_FP64, PREC_MIXED, Engine, metrop1_batch
from tests.gpu_helpers import ENGINE_LIB

pytestmark = pytest.mark.gpu

# (system, walkers, precision): electron counts 10-22, one multideterminant
# system, one all-electron system, one population on the mixed path
CASES = [("small", 5, PREC_FP64), ("c22", 3, PREC_FP64), ("small_md", 4, PREC_FP64), ("small", 6, PREC_MIXED),
         ("small_ae", 2, PREC_FP64)]


def _engines():
    from champ_amd import system as S
    out = []
    for k, (name, nw, prec) in enumerate(CASES):
        sy = S.config(name)
        e = Engine(sy, nw, precision=prec, lib_path=ENGINE_LIB)
        e.set_walkers(sy.initial_walkers(nw, seed=11 + k))
        e.seed_rng(S.seed_from_string())
        e.hpsi()
        out.append(e)
    return out


def test_metrop1_batch_matches_separate_runs(require_gpu):
    batch, single = _engines(), _engines()
    b1 = metrop1_batch(batch, 2, tau=0.3)
    hb = [e.hpsi() for e in batch]
    b2 = metrop1_batch(batch, 1, tau=0.3)
    s1 = [e.metrop1(2, tau=0.3) for e in single]
    hs = [e.hpsi() for e in single]
    s2 = [e.metrop1(1, tau=0.3) for e in single]
    for m, (name, nw, prec) in enumerate(CASES):
        print(f"[batch {name} W={nw} prec={prec}] accepted {b1[m]['accepted']} {s1[m]['accepted']}", flush=True)
        for sa, sb in ((b1[m], s1[m]), (b2[m], s2[m])):
            assert sa["accepted"] == sb["accepted"] and sa["proposed"] == sb["proposed"]
            assert sa["nflagged"] == sb["nflagged"]
            for k in ("esum", "esum2"):
                assert abs(sa[k] - sb[k]) <= 1e-12 * max(1.0, abs(sb[k])), (name, k)
        assert np.array_equal(batch[m].get_walkers(), single[m].get_walkers()), name
        assert np.array_equal(hb[m]["energy"], hs[m]["energy"]), name
        assert np.array_equal(batch[m].scalars()["eold"], single[m].scalars()["eold"]), name
        assert np.array_equal(batch[m].get_rng(), single[m].get_rng()), name


def test_metrop1_batch_rejects_repeated_context(require_gpu):
    e = _engines()[0]
    with pytest.raises(RuntimeError, match="twice"):
        metrop1_batch([e, e], 1, tau=0.3)


def test_metrop1_batch_heterogeneous_molecules(require_gpu):
    """Three molecules of the heterogeneous batch (20, 96 and 300 electrons;
    the last with its Slater inverse in a global work matrix), FP64 and mixed:
    batched sweeps equal each molecule's own sweeps bit for bit."""
    from champ_amd import system as S
    cases = [("het20", 8, PREC_MIXED), ("het96", 4, PREC_FP64), ("het300", 2, PREC_FP64)]

    def make():
        out = []
        for k, (name, nw, prec) in enumerate(cases):
            sy = S.config(name)
            e = Engine(sy, nw, precision=prec, lib_path=ENGINE_LIB)
            e.set_walkers(sy.initial_walkers(nw, seed=21 + k))
            e.seed_rng(S.seed_from_string())
            e.hpsi()
            out.append(e)
        return out

    batch, single = make(), make()
    sb = metrop1_batch(batch, 1, tau=0.3)
    ss = [e.metrop1(1, tau=0.3) for e in single]
    for m, (name, nw, prec) in enumerate(cases):
        print(f"[batch {name} W={nw} prec={prec}] accepted {sb[m]['accepted']} {ss[m]['accepted']}", flush=True)
        assert sb[m]["accepted"] == ss[m]["accepted"] and sb[m]["nflagged"] == ss[m]["nflagged"]
        assert abs(sb[m]["esum"] - ss[m]["esum"]) <= 1e-12 * max(1.0, abs(ss[m]["esum"])), name
        assert np.array_equal(batch[m].get_walkers(), single[m].get_walkers()), name
        assert np.array_equal(batch[m].scalars()["eold"], single[m].scalars()["eold"]), name
