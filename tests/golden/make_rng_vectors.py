"""Write tests/golden/rng_vectors.json from CHAMP's own xoshiro256** C file.

The reference's src/vmc/xoroshiro256starstar.c is compiled from the CHAMP tree
by `make -C oracle ref` into oracle/_ref/libxoshiro_ref.so (never copied).
This script drives that library through the same calls CHAMP makes
(xoroshiro_set_state with the 8 x i4 seed, next_uniform_double, jump) and
stores the outputs as a fixture.
"""
import ctypes as C
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from champ_amd.system import seed_from_string  # noqa: E402
from oracle import ffi  # noqa: E402


def main():
    L = ffi.ref_rng_lib()
    if L is None:
        raise SystemExit("oracle/_ref/libxoshiro_ref.so not built (make -C oracle ref)")
    seed = seed_from_string()
    out = {"seed_string": "18123437465549275475255231234865", "seed_i4": [int(v) for v in seed], "streams": []}
    st = (C.c_uint32 * 8)(*[int(v) & 0xFFFFFFFF for v in seed])
    L.xoroshiro_set_state(st)
    for rank in range(4):
        g = (C.c_uint32 * 8)()
        L.xoroshiro_get_state(g)
        state = [int(g[2 * i]) | (int(g[2 * i + 1]) << 32) for i in range(4)]
        # sample the stream without disturbing the jump chain
        L.xoroshiro_set_state(g)
        nexts = [int(L.next()) for _ in range(8)]
        L.xoroshiro_set_state(g)
        unis = [float(L.next_uniform_double()) for _ in range(8)]
        out["streams"].append({"rank": rank, "state": [str(v) for v in state],
                               "next": [str(v) for v in nexts], "uniform": [repr(u) for u in unis]})
        L.xoroshiro_set_state(g)
        L.jump()
    with open(os.path.join(HERE, "rng_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote rng_vectors.json")


if __name__ == "__main__":
    main()
