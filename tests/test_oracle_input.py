"""Oracle pins for the input path of the non-interactive protocol.

- rlwe_key_switch (src/rgsw/runtime.rs:1001-1063) against a closed form over
  Python ints: with digits d_r of -a (the pinned decomposer restatement) and
  the coefficient-domain key rows A_r, B_r, the output is
  (sum_r d_r * A_r, sum_r d_r * B_r + b) in Z_q[X]/(X^N + 1) -- a canonical
  pair, so this fixes every output word.
- sample_extract (src/pbs.rs:423-449): the LWE decrypts to coefficient k of
  the RLWE's phase b - a*s for every k (Python big-int product).
- the whole path: bits encrypted under a party's u, switched to the ideal s
  and extracted decrypt to the bits; before the switch they decrypt under u.
"""
import numpy as np
import pytest

from conftest import keys_for, oracle_for, rng, uniform_mod
from test_oracle_decomposer import py_digits
from test_oracle_ntt import negacyclic_bigint


def test_rlwe_key_switch_closed_form():
    o = oracle_for("NI_2P")
    q, N, d = o.Q, o.N, o.p.uitos_d
    r = rng(300)
    rlwe = uniform_mod(r, q, (2, N))
    rlwe[0, :3] = [0, 1, q - 1]  # -0 is q: decomposes like 0
    ksk = uniform_mod(r, q, (2 * d, N))
    got = o.rlwe_key_switch(rlwe, ksk, o.shoup_rows(ksk))
    coef = o.ntt_rows("backward", ksk)  # rows in the coefficient domain
    digits = [py_digits(q, o.p.uitos_logb, d, (q - int(v)) % (1 << 64) if int(v) else q)
              for v in rlwe[0]]
    want_a = [0] * N
    want_b = [int(v) for v in rlwe[1]]
    for row in range(d):
        dr = [dg[row] % q for dg in digits]
        pa = negacyclic_bigint(dr, [int(v) for v in coef[row]], q)
        pb = negacyclic_bigint(dr, [int(v) for v in coef[d + row]], q)
        want_a = [(x + y) % q for x, y in zip(want_a, pa)]
        want_b = [(x + y) % q for x, y in zip(want_b, pb)]
    assert got[0].tolist() == want_a
    assert got[1].tolist() == want_b


@pytest.mark.parametrize("k", [0, 1, 777, 2047])
def test_sample_extract_decrypts_to_the_coefficient(k):
    o = oracle_for("NI_2P")
    q, N = o.Q, o.N
    r = rng(301 + k)
    rlwe = uniform_mod(r, q, (2, N))
    s = r.integers(-1, 2, size=N).astype(np.int32)
    lwe = o.sample_extract(rlwe, k)
    phase = negacyclic_bigint([int(v) for v in rlwe[0]], [int(v) % q for v in s], q)
    want = (int(rlwe[1][k]) - phase[k]) % q
    assert o.decrypt_raw(s, lwe) == want


@pytest.mark.parametrize("name", ["NI_2P", "NI_8P"])
def test_input_path_decrypts(name):
    o, s, z, key = keys_for(name)
    u = o.keygen_user_secret(17)
    assert sorted(set(u.tolist())) == [-1, 0, 1] and int(np.count_nonzero(u)) == o.N // 2
    ksk = o.keygen_rlwe_ksk(18, u, s)
    ksk_s = o.shoup_rows(ksk)
    r = rng(302)
    bits = r.integers(0, 2, size=o.N + 37).astype(np.uint8)  # two ciphertexts, one partial
    cts = o.encrypt_rlwe_bools(u, bits, 19)
    assert cts.shape == (2, 2, o.N)
    for t in (0, 5, o.N - 1, o.N, o.N + 36):
        assert o.decrypt_bool(u, o.sample_extract(cts[t // o.N], t % o.N)) == bits[t]
    switched = [o.rlwe_key_switch(ct, ksk, ksk_s) for ct in cts]
    got = [o.decrypt_bool(s, o.sample_extract(switched[t // o.N], t % o.N))
           for t in range(len(bits))]
    assert got == bits.tolist()
