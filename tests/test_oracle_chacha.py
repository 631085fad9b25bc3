"""Pins of the oracle's ChaCha generator (the source of the NTT root psi).

* The 20-round block function against ChaCha20 keystreams produced by
  OpenSSL 3.0.2 (`openssl enc -chacha20`, zero key/nonce) and RFC 7539's
  section 2.4.2 block (key 00..1f, counter 1, nonce 0:0x4a:0).
* The 8-round stream (rand_chacha's ChaCha8Rng layout, 64-bit counter and
  64-bit stream id) against an independent pure-Python implementation.
* psi = the first accepted (next_u64 % Q)^((Q-1)/2N) (src/ntt.rs:283-307),
  recomputed here from the Python stream, and psi^N == -1 (mod Q).
"""
import math

import numpy as np

from conftest import oracle_for
from oracle.oracle import lib, seed_bytes

OPENSSL_ZERO_CTR0 = [0xade0b876, 0x903df1a0, 0xe56a5d40, 0x28bd8653, 0xb819d2bd, 0x1aed8da0,
                     0xccef36a8, 0xc70d778b, 0x7c5941da, 0x8d485751, 0x3fe02477, 0x374ad8b8,
                     0xf4b8436a, 0x1ca11815, 0x69b687c3, 0x8665eeb2]
OPENSSL_ZERO_CTR1 = [0xbee7079f, 0x7a385155, 0x7c97ba98, 0x0d082d73, 0xa0290fcb, 0x6965e348,
                     0x3e53c612, 0xed7aee32, 0x7621b729, 0x434ee69c, 0xb03371d5, 0xd539d874,
                     0x281fed31, 0x45fb0a51, 0x1f0ae1ac, 0x6f4d794b]
RFC7539_242 = [0xf3514f22, 0xe1d91b40, 0x6f27de2f, 0xed1d63b8, 0x821f138c, 0xe2062c3d,
               0xecca4f7e, 0x78cff39e, 0xa30a3b8a, 0x920a6072, 0xcd7479b5, 0x34932bed,
               0x40ba4c79, 0xcd343ec6, 0x4c2c21ea, 0xb7417df0]

M32 = 0xFFFFFFFF


def c_block(key, counter, stream, rounds):
    out = np.zeros(16, dtype=np.uint32)
    lib().pzo_chacha_block(np.frombuffer(bytes(key), dtype=np.uint8).copy(), counter, stream,
                           rounds, out)
    return [int(x) for x in out]


def py_block(key, counter, stream, rounds):
    def rotl(v, c):
        return ((v << c) & M32) | (v >> (32 - c))

    def qr(x, a, b, c, d):
        x[a] = (x[a] + x[b]) & M32; x[d] = rotl(x[d] ^ x[a], 16)
        x[c] = (x[c] + x[d]) & M32; x[b] = rotl(x[b] ^ x[c], 12)
        x[a] = (x[a] + x[b]) & M32; x[d] = rotl(x[d] ^ x[a], 8)
        x[c] = (x[c] + x[d]) & M32; x[b] = rotl(x[b] ^ x[c], 7)

    s = [0x61707865, 0x3320646e, 0x79622d32, 0x6b206574]
    s += [int.from_bytes(bytes(key[4 * i:4 * i + 4]), "little") for i in range(8)]
    s += [counter & M32, counter >> 32, stream & M32, stream >> 32]
    x = list(s)
    for _ in range(rounds // 2):
        qr(x, 0, 4, 8, 12); qr(x, 1, 5, 9, 13); qr(x, 2, 6, 10, 14); qr(x, 3, 7, 11, 15)
        qr(x, 0, 5, 10, 15); qr(x, 1, 6, 11, 12); qr(x, 2, 7, 8, 13); qr(x, 3, 4, 9, 14)
    return [(x[i] + s[i]) & M32 for i in range(16)]


def test_chacha20_openssl_kat():
    z = [0] * 32
    assert c_block(z, 0, 0, 20) == OPENSSL_ZERO_CTR0
    assert c_block(z, 1, 0, 20) == OPENSSL_ZERO_CTR1


def test_chacha20_rfc7539_block():
    key = list(range(32))
    assert c_block(key, 1, 0x4a000000, 20) == RFC7539_242


def test_chacha8_matches_python():
    for key in ([0] * 32, list(range(32)), [0xA5] * 32):
        for ctr in (0, 1, 7, 2 ** 32 + 3):
            assert c_block(key, ctr, 0, 8) == py_block(key, ctr, 0, 8)
            assert c_block(key, ctr, 5, 20) == py_block(key, ctr, 5, 20)


def py_stream_u64(key, count):
    words = []
    ctr = 0
    while len(words) < 2 * count:
        words += py_block(key, ctr, 0, 8)
        ctr += 1
    return [words[2 * i] | (words[2 * i + 1] << 32) for i in range(count)]


def test_rng_stream_order():
    L = lib()
    seed = seed_bytes(0)
    r = L.pzo_rng_new(seed)
    got = [L.pzo_rng_next_u64(r) for _ in range(20)]
    L.pzo_rng_free(r)
    assert got == py_stream_u64([0] * 32, 20)


def psi_from_python(q, two_n):
    t = (q - 1) // two_n
    for v in py_stream_u64([0] * 32, 100):
        omega = pow(v % q, t, q)
        if pow(omega, two_n >> 1, q) == 1:
            continue
        return omega
    raise AssertionError("no root")


def test_psi_is_chacha8_seed0_root():
    for name in ("NI_2P", "SP_TEST"):
        o = oracle_for(name)
        q, n = o.Q, o.N
        psi = o.psi()
        assert psi == psi_from_python(q, 2 * n)
        assert pow(psi, n, q) == q - 1
        assert pow(psi, 2 * n, q) == 1


def test_uniform_sampler_range_and_balance():
    L = lib()
    r = L.pzo_rng_new(seed_bytes(3))
    q = 18014398509404161
    xs = np.array([L.pzo_rng_uniform(r, q) for _ in range(20000)], dtype=np.float64)
    small = [L.pzo_rng_uniform(r, 7) for _ in range(7000)]
    L.pzo_rng_free(r)
    assert xs.max() < q
    assert abs(xs.mean() / q - 0.5) < 0.01
    counts = np.bincount(small, minlength=7)
    assert counts.min() > 850 and counts.max() < 1150


def test_noise_samplers():
    L = lib()
    r = L.pzo_rng_new(seed_bytes(4))
    g = np.array([L.pzo_rng_gauss_round(r) for _ in range(40000)], dtype=np.float64)
    t = np.array([L.pzo_rng_gauss_trunc(r) for _ in range(40000)], dtype=np.float64)
    L.pzo_rng_free(r)
    assert abs(g.mean()) < 0.06 and abs(g.std() - expected_std(0.5)) < 0.05
    assert abs(t.mean()) < 0.06 and abs(t.std() - expected_std(1.0)) < 0.05


def expected_std(off, sigma=3.19):
    """std of |N(0, sigma^2)| rounded (off = 0.5) or truncated (off = 1)."""
    phi = lambda x: 0.5 * (1 + math.erf(x / (sigma * math.sqrt(2))))  # noqa: E731
    var = sum(k * k * 2 * (phi(k + off) - phi(k + off - 1)) for k in range(1, 60))
    return math.sqrt(var)


def test_ternary_hamming_weight():
    L = lib()
    r = L.pzo_rng_new(seed_bytes(5))
    out = np.zeros(2048, dtype=np.int32)
    L.pzo_ternary_hw(r, out, 2048, 1024)
    L.pzo_rng_free(r)
    assert set(np.unique(out)) <= {-1, 0, 1}
    assert int(np.count_nonzero(out)) == 1024
