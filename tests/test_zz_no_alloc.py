"""champ_gpu_basis_fns_noalloc: the allocation-free variant of basis_fns
(no counterpart in the reference), staged in the library's static device
scratch.  Checked against champ_gpu_basis_fns on two calls in a row, the
second smaller, so that a stale value left in the shared scratch by the
first call would show; and its size limit refused on the host."""
import numpy as np
import pytest

from champ_amd import system as S
from champ_amd.engine import Engine
from tests.gpu_helpers import ENGINE_LIB

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["c22", "c160"])
def test_basis_fns_noalloc_matches_basis_fns(require_gpu, name):
    sy = S.config(name)
    e = Engine(sy, 2, lib_path=ENGINE_LIB)
    rng = np.random.default_rng(5)
    limit = ((2 << 20) - 32) // (3 + 5 * sy.nbasis)  # 16 MiB of doubles, values 256-byte aligned
    for npt in (min(limit, 400), 37):
        pos = sy.cent[rng.integers(0, sy.ncent, npt)] + rng.normal(0.0, 1.5, (npt, 3))
        print(f"[noalloc {name}] npt={npt} nbasis={sy.nbasis}", flush=True)
        ref = e.basis_fns(pos)
        got = e.basis_fns_noalloc(pos)
        assert np.array_equal(got, ref)
        assert np.count_nonzero(ref) > 0
    with pytest.raises(RuntimeError, match="16 MiB"):
        e.basis_fns_noalloc(np.zeros((limit + 1, 3)))
    # the largest accepted size works and still matches
    pos = sy.cent[rng.integers(0, sy.ncent, limit)] + rng.normal(0.0, 1.5, (limit, 3))
    assert np.array_equal(e.basis_fns_noalloc(pos), e.basis_fns(pos))
    e.close()
