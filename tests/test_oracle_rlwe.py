"""Pins of the oracle's RLWE operators (src/rgsw/runtime.rs).

* rlwe_by_rgsw_shoup == rlwe_by_rgsw (Barrett twin) bit for bit, trivial and
  not (the reference's invariant, src/rgsw/mod.rs:650).
* rlwe_auto_shoup == rlwe_auto bit for bit (src/rgsw/mod.rs:790).
* Decryption: RLWE(m) x RGSW(X^{emb*z_i}) decrypts to m * X^{emb*z_i} (the
  negacyclic product, computed here with Python ints) and auto_k decrypts to
  m(X^k), both within a small noise bound.
* The automorphism map against Python (src/rgsw/keygen.rs:16-41).
"""
import numpy as np
import pytest

from conftest import keys_for, rng
from test_oracle_ntt import negacyclic_bigint


def rlwe_encrypt(o, s, m, seed):
    """RLWE(m) = [a, a*s + e + m] under the ternary secret s (Python ints)."""
    r = rng(seed)
    q, n = o.Q, o.N
    a = r.integers(0, q, size=n, dtype=np.uint64)
    e = r.integers(-3, 4, size=n)
    s_mod = np.array([int(v) % q for v in s], dtype=np.uint64)
    as_ = negacyclic_bigint(a, s_mod, q)
    b = [(as_[i] + int(e[i]) + int(m[i])) % q for i in range(n)]
    return np.stack([a, np.array(b, dtype=np.uint64)])


def rlwe_decrypt(o, s, ct):
    q = o.Q
    s_mod = np.array([int(v) % q for v in s], dtype=np.uint64)
    as_ = negacyclic_bigint(ct[0], s_mod, q)
    return [(int(ct[1][i]) - as_[i]) % q for i in range(o.N)]


def centered(v, q):
    return v - q if v > q // 2 else v


def monomial_times(m, k, q):
    """m(X) * X^k mod X^N + 1 (k may be negative)."""
    n = len(m)
    out = [0] * n
    k %= 2 * n
    for i, c in enumerate(m):
        t = i + k
        sgn = 1
        if t >= n:
            t -= n
            sgn = -sgn
        if t >= n:
            t -= n
            sgn = -sgn
        out[t] = (out[t] + sgn * int(c)) % q
    return out


@pytest.mark.parametrize("name", ["NI_2P", "SP_TEST"])
def test_ext_product_shoup_equals_barrett_and_decrypts(name):
    o, s, z, key = keys_for(name)
    q, n = o.Q, o.N
    r = rng(50)
    emb = (2 * n) // o.p.br_q
    delta = q // 8
    for trial, idx in enumerate((0, 7, o.n - 1)):
        bits = r.integers(0, 2, size=n)
        m = [int(b) * delta for b in bits]
        ct = rlwe_encrypt(o, s, m, 60 + trial)
        trivial = trial == 1
        if trivial:
            ct = np.stack([np.zeros(n, dtype=np.uint64), np.array(m, dtype=np.uint64)])
        rg, rgs = key.rgsw_ct(idx)
        a = o.rlwe_by_rgsw_shoup(ct, rg, rgs, trivial)
        b = o.rlwe_by_rgsw(ct, rg, trivial)
        assert np.array_equal(a, b)
        dec = rlwe_decrypt(o, s, a)
        want = monomial_times(m, emb * int(z[idx]), q)
        noise = max(abs(centered((d - w) % q, q)) for d, w in zip(dec, want))
        assert noise < q // 64, noise


@pytest.mark.parametrize("name", ["NI_2P", "SP_TEST"])
def test_auto_shoup_equals_barrett_and_decrypts(name):
    o, s, z, key = keys_for(name)
    q, n = o.Q, o.N
    r = rng(51)
    delta = q // 8
    for k in range(o.p.w + 1):
        bits = r.integers(0, 2, size=n)
        m = [int(b) * delta for b in bits]
        ct = rlwe_encrypt(o, s, m, 80 + k)
        kt, kts = key.auto_ct(k)
        a = o.rlwe_auto_shoup(ct, kt, kts, k, False)
        b = o.rlwe_auto(ct, kt, k, False)
        assert np.array_equal(a, b)
        idx, sg = o.auto_map(k)
        want = [0] * n
        for i in range(n):
            want[idx[i]] = m[i] % q if sg[i] else (-m[i]) % q
        dec = rlwe_decrypt(o, s, a)
        noise = max(abs(centered((d - w) % q, q)) for d, w in zip(dec, want))
        assert noise < q // 64, (k, noise)
        # trivial path: only b is permuted (and zero maps to q under neg)
        tr = np.stack([np.zeros(n, dtype=np.uint64), np.array(m, dtype=np.uint64)])
        ta = o.rlwe_auto_shoup(tr, kt, kts, k, True)
        tb = o.rlwe_auto(tr, kt, k, True)
        assert np.array_equal(ta, tb)
        assert [int(v) % q for v in ta[1]] == want


def test_auto_map_matches_python():
    o, s, z, key = keys_for("NI_2P")
    n = o.N
    g, brq = o.p.g, o.p.br_q
    for k in range(o.p.w + 1):
        el = -g if k == 0 else pow(g, k) % brq
        kk = el % (2 * n)
        idx, sg = o.auto_map(k)
        for i in range(0, n, 13):
            t = (i * kk) % (2 * n)
            assert (int(idx[i]), bool(sg[i])) == ((t - n, False) if t >= n else (t, True))
