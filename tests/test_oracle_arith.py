"""Pins of the oracle's scalar arithmetic against exact integer arithmetic.

ShoupMul (src/utils.rs:35-47), Barrett mul_mod_fast / _lazy
(src/backend/modulus_u64.rs:83-130), mod_exponent (src/utils.rs:145-165) and
shoup_matrix_fma (src/backend/modulus_u64.rs:231-264, with its reference test
at :285-336) are checked against Python integers, including their output
ranges.
"""
import numpy as np

from conftest import rng
from oracle.oracle import lib

Q = 18014398509404161
PRIMES = [Q, 268369921, 36028797017456641, 1152921504606748673]


def test_q_facts():
    # Q = 2^54 - 77823, Q = 1 mod 4096, not 1 mod 8192, (double)Q = Q - 1
    assert Q == 2 ** 54 - 77823
    assert Q % 4096 == 1 and Q % 8192 != 1
    assert int(float(Q)) == Q - 1
    assert pow(3, Q - 1, Q) == 1 and pow(5, Q - 1, Q) == 1


def test_shoup_repr_and_mul():
    L = lib()
    r = rng(1)
    for q in PRIMES:
        ws = r.integers(0, q, size=200, dtype=np.uint64)
        ys = [int(v) for v in r.integers(0, 2 ** 63, size=200, dtype=np.uint64)]
        ys += [0, 1, 2 ** 64 - 1, 4 * q - 1]
        for w in [int(v) for v in ws] + [0, 1, q - 1]:
            wsh = L.pzo_shoup_repr(w, q)
            assert wsh == (w << 64) // q
            for y in ys[:20]:
                t = L.pzo_shoup_mul(y, w, wsh, q)
                assert t < 2 * q
                assert t % q == (y * w) % q


def test_barrett_mul_mod_fast():
    L = lib()
    r = rng(2)
    for q in PRIMES[:3]:
        a = [int(v) for v in r.integers(0, 2 * q, size=3000, dtype=np.uint64)]
        b = [int(v) for v in r.integers(0, 2 * q, size=3000, dtype=np.uint64)]
        a += [0, 1, 2 * q - 1, q, q - 1]
        b += [2 * q - 1, q - 1, 2 * q - 1, q, 1]
        for x, y in zip(a, b):
            assert L.pzo_mul_mod_fast(q, x, y) == (x * y) % q
            lz = L.pzo_mul_mod_fast_lazy(q, x, y)
            assert lz < 2 * q and lz % q == (x * y) % q


def test_mod_exponent():
    L = lib()
    r = rng(3)
    for q in PRIMES:
        for _ in range(50):
            a = int(r.integers(0, q))
            e = int(r.integers(0, 2 ** 62))
            assert L.pzo_mod_exponent(a, e, q) == pow(a, e, q)


def test_shoup_matrix_fma_reference_case():
    """src/backend/modulus_u64.rs:285-336: prime 36028797017456641, N=8, d=2."""
    L = lib()
    r = rng(4)
    for q, n, d in ((36028797017456641, 8, 2), (Q, 2048, 4), (268369921, 512, 3)):
        a = r.integers(0, q, size=(d, n), dtype=np.uint64)
        a_s = np.array([[(int(v) << 64) // q for v in row] for row in a], dtype=np.uint64)
        b = r.integers(0, 2 ** 64 - 1, size=(d, n), dtype=np.uint64)
        out0 = r.integers(0, 2 * q, size=n, dtype=np.uint64)
        out = out0.copy()
        L.pzo_shoup_matrix_fma(q, out, a.reshape(-1), a_s.reshape(-1), b.reshape(-1), d, n)
        for j in range(n):
            want = (int(out0[j]) + sum(int(a[k, j]) * int(b[k, j]) for k in range(d))) % q
            assert int(out[j]) < 2 * q
            assert int(out[j]) % q == want
