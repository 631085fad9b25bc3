"""Pins of the oracle's gadget decomposer (src/decomposer.rs).

* decompose_iter == decompose_to_vec (the reference's own test, :326-370),
  recomposition within +-1 of the rounded value, for prime and power-of-two q
  of 37 and 55 bits with logB 11, d 3.
* Digit bounds: every digit is k or q + k - B with |signed digit| <= B/2.
* Exact signed digits against a Python big-integer restatement of the
  carry rule, and the release-build wrap of v == q to the digits of 0.
* ignore_bits of the parameter sets (37 / 30 / 3).
"""
import numpy as np
import pytest

from conftest import oracle_for, rng
from oracle.oracle import lib
from test_oracle_ntt import find_prime


def decompose(q, logb, d, vals, fn="iter"):
    L = lib()
    v = np.ascontiguousarray(vals, dtype=np.uint64)
    out = np.zeros(len(v) * d, dtype=np.uint64)
    f = L.pzo_decompose_iter if fn == "iter" else L.pzo_decompose_to_vec
    f(q, logb, d, v, out, len(v))
    return out.reshape(len(v), d)


def round_value(v, ib):
    if ib == 0:
        return v
    return ((v >> ib) + ((v & ((1 << ib) - 1)) >> (ib - 1))) & ((1 << 64) - 1)


def py_digits(q, logb, d, v):
    """Signed digits of the reference's decomposition, as Python ints."""
    logq = q.bit_length() - 1 if q & (q - 1) == 0 else q.bit_length()
    ib = logq - logb * d
    M64 = (1 << 64) - 1
    if v >= q >> 1:
        v = (~(q - v) + 1) & M64
    v = round_value(v, ib)
    B = 1 << logb
    out = []
    for _ in range(d):
        k = v & (B - 1)
        v = (v - k) >> logb
        carry = k > B // 2 or (k == B // 2 and v & 1)
        if carry:
            v += 1
            out.append(k - B)
        else:
            out.append(k)
    return out


@pytest.mark.parametrize("logq", [37, 55])
@pytest.mark.parametrize("prime", [True, False])
def test_reference_decomposition_case(logq, prime):
    q = find_prime(logq, 2 * 2048) if prime else 1 << logq
    logb, d = 11, 3
    L = lib()
    ib = L.pzo_decomposer_ignore_bits(q, logb, d)
    r = rng(20 + logq)
    vals = r.integers(0, q, size=20000, dtype=np.uint64)
    a = decompose(q, logb, d, vals, "iter")
    b = decompose(q, logb, d, vals, "vec")
    assert np.array_equal(a, b)
    for v, limbs in zip(vals[:3000], a[:3000]):
        back = L.pzo_recompose(q, logb, d, np.ascontiguousarray(limbs))
        assert abs(int(round_value(int(v), ib)) - int(round_value(back, ib))) <= 1 or \
            abs(int(round_value(int(v), ib)) - int(round_value(back, ib))) % q <= 1


@pytest.mark.parametrize("name", ["NI_2P", "SP_TEST", "NI_8P"])
def test_parameter_decomposers_exact(name):
    o = oracle_for(name)
    p = o.p
    r = rng(30)
    for q, logb, d in ((p.rlwe_q, p.rlrg_logb, p.rlrg_da), (p.rlwe_q, p.auto_logb, p.auto_d),
                       (p.lwe_q, p.lwe_logb, p.lwe_d)):
        vals = [int(v) for v in r.integers(0, q, size=3000, dtype=np.uint64)]
        vals += [0, 1, q - 1, q // 2, q // 2 - 1, q // 2 + 1]
        got = decompose(q, logb, d, vals)
        B = 1 << logb
        for v, row in zip(vals, got):
            want = py_digits(q, logb, d, v)
            signed = [int(x) if int(x) < q // 2 else int(x) - q for x in row]
            assert signed == want
            assert all(-B // 2 <= s <= B // 2 for s in signed)
            assert all(int(x) < q for x in row)


def test_wrap_of_q_decomposes_like_zero():
    """moddown can return q_ks itself (src/pbs.rs:128-132); the release build
    wraps !(q - v) + 1 to 0 (src/decomposer.rs:219-220)."""
    o = oracle_for("NI_2P")
    p = o.p
    assert np.array_equal(decompose(p.lwe_q, p.lwe_logb, p.lwe_d, [p.lwe_q]),
                          decompose(p.lwe_q, p.lwe_logb, p.lwe_d, [0]))
    assert np.array_equal(decompose(p.rlwe_q, p.rlrg_logb, 1, [p.rlwe_q]),
                          decompose(p.rlwe_q, p.rlrg_logb, 1, [0]))


def test_ignore_bits():
    L = lib()
    o = oracle_for("NI_2P")
    p = o.p
    assert L.pzo_decomposer_ignore_bits(p.rlwe_q, p.rlrg_logb, 1) == 37
    assert L.pzo_decomposer_ignore_bits(p.rlwe_q, p.auto_logb, 1) == 30
    assert L.pzo_decomposer_ignore_bits(p.lwe_q, p.lwe_logb, p.lwe_d) == 3
    p8 = oracle_for("NI_8P").p
    assert L.pzo_decomposer_ignore_bits(p8.lwe_q, p8.lwe_logb, p8.lwe_d) == 3
