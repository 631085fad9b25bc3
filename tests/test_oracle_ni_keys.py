"""Oracle pins for the non-interactive automorphism-key and LWE-KSK shares,
their sums and the expansion of a seeded server key
(src/bool/evaluator.rs:1574-1612, :1952-2046; src/bool/keys.rs:851-945;
src/rgsw/keygen.rs:306-430; src/lwe.rs:46-101).

Every check works from the definitions, not from the oracle's code:
- each automorphism-key row of the expanded key, back in the coefficient
  domain, decrypts under the ideal secret s = sum_j s_j to
  beta_j * (-s(X^k)) with |e| within the summed share noise, where s(X^k) is
  built here from the map i -> i*k mod 2N (a wrong sign, element, gadget
  power or row fails);
- every LWE-KSK row decrypts under z = sum_j z_j to beta_j * (-s_i) mod q_ks
  within the same bound;
- the a parts are the CRS streams: the first draws of
  ChaCha8(puncture(puncture(crs, 1), 2)) (automorphism keys, UniformInt mod
  Q) and of puncture(puncture(crs, 1), 3) (LWE KSK, mod q_ks), from the
  independent Python block function and rand's widening-multiply rule;
- a server key whose automorphism keys and LWE KSK come from the parties'
  shares bootstraps every gate correctly under the ideal secret.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import oracle_for
from test_oracle_chacha import py_stream_u64
from test_oracle_ni import centred, parties, phase

CRS = bytes(range(40, 72))
BOUND = 6 * 4  # per party: rounded Gaussians of sigma 3.19 stay within 6 sigma


def shares(o, total):
    ps = parties(o, total)
    sh = [o.ni_share_auto_ksk(CRS, 5000 + j, ps[j][0], ps[j][1]) for j in range(total)]
    s = sum(p[0].astype(np.int64) for p in ps)
    z = sum(p[1].astype(np.int64) for p in ps)
    auto_b, ksk_b = o.ni_aggregate_auto_ksk(np.stack([x[0] for x in sh]),
                                            np.stack([x[1] for x in sh]))
    return ps, sh, s, z, auto_b, ksk_b


def coeff_rgsw_zero(o):
    rows = 2 * (o.p.rlrg_da + o.p.rlrg_db)
    return np.zeros((o.p.lwe_n, rows, o.N), dtype=np.uint64)


def auto_poly(o, s, k_el):
    """s(X^k_el) in Z_q[X]/(X^N + 1), from the map i -> i*k mod 2N."""
    q, N = o.Q, o.N
    out = [0] * N
    for i, c in enumerate(s):
        t = (i * k_el) % (2 * N)
        v = int(c) % q
        if t >= N:
            t -= N
            v = (-v) % q
        out[t] = v
    return out


def uniform_stream(seed, rng_range, count):
    """rand 0.8.5 UniformInt<u64>::new_inclusive(0, range - 1) draws."""
    zone = (1 << 64) - 1 - ((1 << 64) % rng_range)
    out, words = [], py_stream_u64(list(seed), 4 * count)
    for v in words:
        m = v * rng_range
        if m & ((1 << 64) - 1) <= zone:
            out.append(m >> 64)
            if len(out) == count:
                return out
    raise AssertionError("stream too short")


def test_sums_are_elementwise():
    o = oracle_for("SP_TEST")
    _, sh, _, _, auto_b, ksk_b = shares(o, 3)
    want_a = sum(x[0].astype(object) for x in sh) % o.Q
    want_k = sum(x[1].astype(object) for x in sh) % o.p.lwe_q
    assert np.array_equal(auto_b.astype(object), want_a)
    assert np.array_equal(ksk_b.astype(object), want_k)


@pytest.mark.parametrize("name,total", [("NI_2P", 2), ("SP_TEST", 3)])
def test_auto_key_rows_decrypt(name, total):
    o = oracle_for(name)
    q, N, d = o.Q, o.N, o.p.auto_d
    _, _, s, _, auto_b, ksk_b = shares(o, total)
    key = o.ni_expand_server_key(CRS, coeff_rgsw_zero(o), auto_b, ksk_b)
    ib = q.bit_length() - o.p.auto_logb * d
    g, br_q = o.p.g, o.p.br_q
    for k in range(o.p.w + 1):
        el = (-g) % (2 * N) if k == 0 else pow(g, k, br_q)
        ct, _ = key.auto_ct(k)
        coef = o.ntt_rows("backward", ct)
        neg = [(-v) % q for v in auto_poly(o, s, el)]
        for j in range(d):
            got = phase(o, coef[j], coef[d + j], s)
            want = [(v << (ib + j * o.p.auto_logb)) % q for v in neg]
            err = max(abs(centred((a - b) % q, q)) for a, b in zip(got, want))
            assert err <= BOUND * total, (k, j, err)


@pytest.mark.parametrize("name,total", [("NI_2P", 2), ("NI_8P", 3)])
def test_lwe_ksk_rows_decrypt(name, total):
    o = oracle_for(name)
    p = o.p
    lq, n, d = p.lwe_q, p.lwe_n, p.lwe_d
    _, _, s, z, auto_b, ksk_b = shares(o, total)
    key = o.ni_expand_server_key(CRS, coeff_rgsw_zero(o), auto_b, ksk_b)
    rows = key.lwe_ksk.reshape(o.N * d, n + 1).astype(np.int64)
    az = (rows[:, 1:] * z[None, :]).sum(axis=1)
    phase_ = (rows[:, 0] - az) % lq
    ib = lq.bit_length() - 1 - p.lwe_logb * d
    beta = np.array([1 << (ib + j * p.lwe_logb) for j in range(d)], dtype=np.int64)
    want = (np.repeat(-s, d) * np.tile(beta, o.N)) % lq
    diff = (phase_ - want) % lq
    diff = np.where(diff > lq // 2, diff - lq, diff)
    assert np.abs(diff).max() <= BOUND * total
    assert np.array_equal(rows[:, 0], ksk_b.astype(np.int64))  # b parts copied


def test_a_parts_are_the_crs_streams():
    o = oracle_for("NI_2P")
    _, _, _, _, auto_b, ksk_b = shares(o, 2)
    key = o.ni_expand_server_key(CRS, coeff_rgsw_zero(o), auto_b, ksk_b)
    seed_auto = o.puncture(o.puncture(CRS, 1), 2).tobytes()
    seed_ksk = o.puncture(o.puncture(CRS, 1), 3).tobytes()
    ct, _ = key.auto_ct(0)
    a0 = o.ntt_rows("backward", ct[:1])[0]
    assert [int(v) for v in a0[:16]] == uniform_stream(seed_auto, o.Q, 16)
    n = o.p.lwe_n
    first = key.lwe_ksk.reshape(-1, n + 1)[:2, 1:].reshape(-1)
    assert [int(v) for v in first[:2 * n]] == uniform_stream(seed_ksk, o.p.lwe_q, 2 * n)


GATE_FN = {0: lambda a, b: 1 - (a & b), 1: lambda a, b: a & b, 2: lambda a, b: a | b,
           3: lambda a, b: 1 - (a | b), 4: lambda a, b: a ^ b, 5: lambda a, b: 1 - (a ^ b)}


@pytest.mark.parametrize("name,gates", [("SP_TEST", range(6)), ("NI_2P", (0, 4))])
def test_key_from_shares_bootstraps(name, gates):
    """Automorphism keys and LWE KSK from 2 parties' shares, RGSWs under the
    same ideal secret: every gate of every input pair decrypts correctly."""
    from oracle.oracle import OracleKey, seed_bytes
    o = oracle_for(name)
    ps, _, s, z, auto_b, ksk_b = shares(o, 2)
    s32, z32 = s.astype(np.int32), z.astype(np.int32)
    base = OracleKey(o.sizes, o.N)
    o.L.pzo_keygen_server(o.ev, seed_bytes(9), s32, z32, C.byref(base.cstruct()))
    rows = 2 * (o.p.rlrg_da + o.p.rlrg_db)
    rgsw = o.ntt_rows("backward", base.rgsw.reshape(-1, o.N)).reshape(o.p.lwe_n, rows, o.N)
    key = o.ni_expand_server_key(CRS, rgsw, auto_b, ksk_b)
    assert np.array_equal(key.rgsw, base.rgsw)
    pairs = [(0, 0), (0, 1), (1, 0), (1, 1)] if name == "SP_TEST" else [(1, 1), (1, 0)]
    for g in gates:
        for i, (a, b) in enumerate(pairs):
            c0 = o.encrypt_bool(s32, a, 100 + 10 * g + i)
            c1 = o.encrypt_bool(s32, b, 200 + 10 * g + i)
            assert o.decrypt_bool(s32, o.gate(key, g, c0, c1)) == GATE_FN[g](a, b), (g, a, b)
