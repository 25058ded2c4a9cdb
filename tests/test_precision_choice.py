"""Why the mixed path's orbital contraction takes FP32 (or FP64) operands and
not BF16: the ratio of a single-electron move, sum_j A^-1[j][e] phi_j(r'),
computed from orbitals contracted with operands rounded to each format
(FP32 accumulation, as the matrix cores would), against FP64, on the c160
molecule (600 AOs, 80 orbitals per spin).  The AO values and the Slater
inverse come from the oracle; BF16 rounding is round-to-nearest-even of the
FP32 value to 8 significand bits.  CPU only."""
import ctypes as C

import numpy as np

from champ_amd import system as S
from oracle import ffi


def _bf16(a):
    a32 = np.asarray(a, dtype=np.float32).copy()
    u = a32.view(np.uint32)
    u[:] = ((u.astype(np.uint64) + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16).astype(np.uint32)
    return a32


def test_bf16_rounding_helper():
    # in [1, 2) the BF16 spacing is 2^-7: ties go to the even neighbour, above a
    # tie rounds up, below it down
    x = np.array([1.0 + 2.0 ** -8, 1.0 + 2.0 ** -7 + 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -10, 1.0 + 2.0 ** -9],
                 dtype=np.float32)
    assert _bf16(x).tolist() == [1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, 1.0]


def test_bf16_operands_break_the_move_ratio():
    sy = S.config("c160")
    cs, L, nb = sy.as_ctypes(), ffi.lib(), sy.nbasis

    def phi_at(p):
        rvec = np.ascontiguousarray(p[None, :] - sy.cent)
        r = np.sqrt((rvec ** 2).sum(1))
        ph, dph, d2 = np.zeros(nb), np.zeros(3 * nb), np.zeros(nb)
        L.orc_basis_fns_point(C.byref(cs), ffi.P(rvec), ffi.P(r), 0, ffi.P(ph), ffi.P(dph), ffi.P(d2))
        return ph

    coef = np.asarray(sy.coef)
    iw = np.asarray(sy.iworbd)[0]
    nup = sy.nup
    x = sy.initial_walkers(2, seed=2)
    e32, e16 = [], []
    for w in range(2):
        orb = np.array([phi_at(x[w, e]) for e in range(nup)]) @ coef.T
        ainv = np.linalg.inv(orb[:, iw[:nup]])
        for e in range(0, nup, 8):
            ph = phi_at(x[w, e] + np.random.default_rng(e).normal(0, 0.3, 3))
            ratio = lambda o: ainv[:, e] @ o[iw[:nup]]
            r64 = ratio(coef @ ph)
            e32.append(abs(ratio((coef.astype(np.float32) @ ph.astype(np.float32)).astype(np.float64)) - r64) / abs(r64))
            e16.append(abs(ratio((_bf16(coef) @ _bf16(ph)).astype(np.float64)) - r64) / abs(r64))
    print(f"[precision] move ratio relative error: FP32 median {np.median(e32):.1e} max {np.max(e32):.1e}; "
          f"BF16 median {np.median(e16):.1e} max {np.max(e16):.1e}")
    # FP32 operands keep the acceptance ratio to about 1e-7 (typically); BF16
    # operands move it by several parts per thousand, and by percent at times,
    # which would bias the sampled distribution
    assert np.median(e32) < 1e-5 and np.median(e16) > 1e-3 and np.max(e16) > 100 * np.max(e32)
