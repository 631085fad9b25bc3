"""Writes tests/golden/golden.npz: small input/output vectors produced by the
oracle (the C restatement of the reference's PBS path).

The reference ships no golden vectors (its tests are invariant-style and draw
from thread_rng), so these fixtures pin the oracle itself against drift and
give the GPU tests fixed expected outputs.  Run:  python tests/golden/gen_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import Oracle  # noqa: E402
from oracle.oracle import lib, seed_bytes  # noqa: E402


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return np.frombuffer(h.digest(), dtype=np.uint8).copy()


def main():
    out = {}
    L = lib()
    r = L.pzo_rng_new(seed_bytes(0))
    out["chacha8_seed0_u64"] = np.array([L.pzo_rng_next_u64(r) for _ in range(16)], dtype=np.uint64)
    L.pzo_rng_free(r)

    for name in ("SP_TEST", "NI_2P"):
        o = Oracle(name)
        g = np.random.default_rng(123)
        out[name + "_psi"] = np.array([o.psi()], dtype=np.uint64)
        x = g.integers(0, o.Q, size=(2, o.N), dtype=np.uint64)
        out[name + "_ntt_in"] = x
        for which in ("forward", "forward_lazy", "backward", "backward_lazy"):
            out[name + "_ntt_" + which] = o.ntt_rows(which, x)
        s, z, key = o.keygen(0, 1, {"SP_TEST": 1, "NI_2P": 2}[name])
        out[name + "_key_digest"] = digest(key.rgsw, key.auto_key, key.lwe_ksk, s, z)
        rl = g.integers(0, o.Q, size=(2, o.N), dtype=np.uint64)
        rg, rgs = key.rgsw_ct(3)
        out[name + "_ext_in"] = rl
        out[name + "_ext_out"] = o.rlwe_by_rgsw_shoup(rl, rg, rgs, False)
        kt, kts = key.auto_ct(2)
        out[name + "_auto_out"] = o.rlwe_auto_shoup(rl, kt, kts, 2, False)
        bits = [(0, 0), (0, 1), (1, 0), (1, 1)]
        c0 = np.stack([o.encrypt_bool(s, a, 500 + i) for i, (a, _) in enumerate(bits)])
        c1 = np.stack([o.encrypt_bool(s, b, 600 + i) for i, (_, b) in enumerate(bits)])
        out[name + "_gate_c0"] = c0
        out[name + "_gate_c1"] = c1
        gates = np.array([0, 1, 4, 5], dtype=np.uint32)  # nand, and, xor, xnor
        out[name + "_gate_ops"] = gates
        out[name + "_gate_out"] = np.stack([o.gate(key, int(gates[i]), c0[i], c1[i])
                                            for i in range(4)])
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), sum(v.nbytes for v in out.values()), "bytes")


if __name__ == "__main__":
    main()
