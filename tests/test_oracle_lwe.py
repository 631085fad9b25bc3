"""Pins of the oracle's LWE-side steps.

* f64 moddown Q -> q_ks (src/pbs.rs:128-132) against Python floats (IEEE
  doubles), at random values, at rounding boundaries and in the region where
  it returns q_ks itself (v >= 18014261070450687 at NI_2P).
* mod_switch_odd (src/pbs.rs:416-420) against its integer closed form
  ((v * br_q) >> log2(q_ks)) | 1 for every v < q_ks.
* lwe_key_switch (src/lwe.rs:14-44) against a Python big-integer dot product
  of the decomposition digits with the KSK rows, and by decryption under the
  LWE secret (small noise).
"""
import numpy as np

from conftest import keys_for, oracle_for, rng
from test_oracle_decomposer import py_digits


def py_moddown(v, lwe_q, Q):
    return int(round_half_away(float(v) * float(lwe_q) / float(Q)))


def round_half_away(x):
    import math
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def test_moddown_matches_ieee_doubles():
    o = oracle_for("NI_2P")
    Q, lq = o.Q, o.p.lwe_q
    r = rng(40)
    vals = [int(v) for v in r.integers(0, Q, size=20000, dtype=np.uint64)]
    # values around the rounding boundaries k + 1/2 of v * 2^16 / Q
    for k in range(0, 65536, 997):
        c = ((2 * k + 1) * Q) // (2 * lq)
        vals += [c - 2, c - 1, c, c + 1, c + 2]
    vals += [Q - 1, Q - 2, 18014261070450687, 18014261070450686, 0, 1]
    for v in vals:
        assert o.moddown(v) == py_moddown(v, lq, Q)


def test_moddown_wrap_region():
    o = oracle_for("NI_2P")
    assert o.moddown(18014261070450687) == 65536
    assert o.moddown(18014261070450686) == 65535
    assert o.moddown(o.Q - 1) == 65536


def test_mod_switch_odd_closed_form():
    from oracle.oracle import lib
    L = lib()
    for lq, bq in ((1 << 16, 1 << 12), (1 << 16, 1 << 11), (1 << 17, 1 << 12), (1 << 16, 1 << 9)):
        shift = lq.bit_length() - 1
        for v in range(0, lq, 7):
            assert L.pzo_mod_switch_odd(float(v), float(lq), float(bq)) == ((v * bq) >> shift) | 1


def test_key_switch_bigint():
    o, s, z, key = keys_for("NI_2P")
    p = o.p
    n1 = p.lwe_n + 1
    ksk = key.lwe_ksk.reshape(o.N * p.lwe_d, n1)
    r = rng(41)
    x = r.integers(0, p.lwe_q, size=o.N + 1, dtype=np.uint64)
    x[5] = p.lwe_q  # moddown's q_ks output decomposes like 0
    got = o.lwe_key_switch(x, ksk)
    acc = np.zeros(n1, dtype=object)
    for i in range(o.N):
        for j, dg in enumerate(py_digits(p.lwe_q, p.lwe_logb, p.lwe_d, int(x[1 + i]))):
            if dg:
                acc += dg * ksk[i * p.lwe_d + j].astype(object)
    acc[0] += int(x[0])
    want = np.array([int(v) % p.lwe_q for v in acc], dtype=np.uint64)
    assert np.array_equal(got, want)


def test_key_switch_decrypts_under_lwe_secret():
    o, s, z, key = keys_for("NI_2P")
    p = o.p
    n1 = p.lwe_n + 1
    ksk = key.lwe_ksk.reshape(o.N * p.lwe_d, n1)
    lq = p.lwe_q
    r = rng(42)
    for m in (0, lq // 4, lq // 2, 3 * lq // 4):
        a = r.integers(0, lq, size=o.N, dtype=np.int64)
        e = int(r.integers(-4, 5))
        b = (int(np.dot(a, s.astype(np.int64))) + m + e) % lq
        x = np.concatenate([[b], a]).astype(np.uint64)
        out = o.lwe_key_switch(x, ksk)
        dec = (int(out[0]) - int(np.dot(out[1:].astype(np.int64), z.astype(np.int64)))) % lq
        err = (dec - m + lq // 2) % lq - lq // 2
        assert abs(err) < lq // 16, err
