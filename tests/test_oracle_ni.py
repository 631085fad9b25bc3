"""Oracle pins for the non-interactive server-key shares and their
aggregation (src/bool/evaluator.rs:480-608, :1222-1565, :1625-1937;
src/multi_party.rs:186-291).

Every check decrypts with the ideal secret over Python ints (Kronecker
products), so a wrong sign, gadget power, row alignment or party order shows
up as a phase far from the expected message:
- puncture(seed, t) is the t-th seed of the ChaCha8 stream, each byte the low
  byte of one stream word (the independent Python block function);
- the leader of index l is the party whose segment [per*j, per*(j+1)) holds l,
  per = ceil(lwe_n / total);
- Ksk(u_j -> s) rows satisfy B_i - A_i*s = beta_i*u_j + e with |e| within the
  summed share noise;
- produced RGSW rows (leader: RLWE x RGSW counts, coefficient domain; others:
  RGSW x RGSW counts, evaluation domain) are RLWE'(-s*X^{emb*z_j[l]}) and
  RLWE'(X^{emb*z_j[l]}) under s with |e| < 2^13, below the smallest gadget
  power 2^14;
- the aggregate is the leader's RGSW times the others' in ascending order
  (the pinned RGSW x RGSW product), and decrypts to X^{emb*(sum_j z_j[l])}.
"""
import numpy as np
import pytest

from conftest import oracle_for
from test_oracle_chacha import py_block
from test_oracle_ntt import negacyclic_bigint

CRS = bytes(range(32))


def phase(o, a_row, b_row, s):
    q = o.Q
    prod = negacyclic_bigint([int(v) for v in a_row], [int(v) % q for v in s], q)
    return [(int(b) - p) % q for b, p in zip(b_row, prod)]


def centred(v, q):
    return v - q if v > q // 2 else v


def mono_times(o, poly, e):
    """poly(X) * X^e in Z_q[X]/(X^N + 1), e in [-N, N)."""
    q, N = o.Q, o.N
    e %= 2 * N
    out = [0] * N
    for i, c in enumerate(poly):
        k = i + e
        sign = 1
        while k >= N:
            k -= N
            sign = -sign
        out[k] = (out[k] + sign * int(c)) % q
    return out


def max_err(o, got, want):
    q = o.Q
    return max(abs(centred((g - w) % q, q)) for g, w in zip(got, want))


def parties(o, total):
    out = []
    for j in range(total):
        s, z = o.keygen_party(1000 + j)
        u = o.keygen_user_secret(2000 + j)
        out.append((s, z, u))
    return out


def test_puncture_is_the_stream_chunk():
    o = oracle_for("NI_2P")
    seed = bytes(range(100, 132))
    words = sum((py_block(seed, c, 0, 8) for c in range(8)), [])
    for t in (1, 2, 3, 4):
        want = bytes(w & 0xFF for w in words[32 * (t - 1):32 * t])  # one word per byte
        assert o.puncture(seed, t).tobytes() == want


@pytest.mark.parametrize("total", [1, 2, 3, 4, 7, 8])
def test_leader_segments(total):
    o = oracle_for("NI_8P")
    n = o.p.lwe_n
    per = -(-n // total)
    for l in range(n):
        j = o.ni_leader(l, total)
        assert per * j <= l < min(per * (j + 1), n)


def uitos_ksks(o, total, ps):
    shares = [o.ni_share_uitos(CRS, 3000 + j, j, total, ps[j][0], ps[j][2]) for j in range(total)]
    ksks = []
    for j in range(total):
        zf = [shares[k][1][j if j < k else j - 1] for k in range(total) if k != j]
        zf = np.array(zf, dtype=np.uint64).reshape(total - 1, o.p.uitos_d, o.N)
        ksks.append(o.ni_uitos_ksk(CRS, total, j, shares[j][0], zf))
    return ksks


@pytest.mark.parametrize("name,total", [("NI_2P", 2), ("NI_4P_HB_FR", 3)])
def test_uitos_ksk_decrypts(name, total):
    o = oracle_for(name)
    q, N, d = o.Q, o.N, o.p.uitos_d
    ps = parties(o, total)
    s = sum(p[0].astype(np.int64) for p in ps)
    ksks = uitos_ksks(o, total, ps)
    ib = 54 - o.p.uitos_logb * d
    for j in (0, total - 1):
        coef = o.ntt_rows("backward", ksks[j])
        u = ps[j][2]
        for i in (0, 1, d // 2, d - 1):
            got = phase(o, coef[i], coef[d + i], s)
            want = [(int(x) * (1 << (ib + i * o.p.uitos_logb))) % q for x in u]
            assert max_err(o, got, want) <= 6 * 4 * total  # summed rounded Gaussians


def check_rgsw(o, rows, s, e_exp, counts, logb, bound):
    """rows [2da+2db][N] coefficient domain: RLWE'(-s X^e), RLWE'(X^e)."""
    q, N = o.Q, o.N
    da, db = counts
    worst = 0
    for k in range(da + db):
        nsm = k < da
        ra, rb = (k, da + k) if nsm else (2 * da + k - da, 2 * da + db + k - da)
        d_part = da if nsm else db
        kk = k if nsm else k - da
        beta = 1 << (54 - logb * d_part + logb * kk)
        base = [(-int(x)) % q for x in s] if nsm else [1] + [0] * (N - 1)
        want = [(v * beta) % q for v in mono_times(o, base, e_exp)]
        err = max_err(o, phase(o, rows[ra], rows[rb], s), want)
        worst = max(worst, err)
        assert err < bound, (k, err)
    return worst


@pytest.mark.parametrize("name", ["NI_2P", "NI_4P_HB_FR"])
def test_produce_decrypts(name):
    o = oracle_for(name)
    p = o.p
    total = 2
    ps = parties(o, total)
    s = sum(x[0].astype(np.int64) for x in ps)
    ksks = uitos_ksks(o, total, ps)
    per = -(-p.lwe_n // total)
    for l in (0, per):  # led by party 0, then by party 1
        shares = [o.ni_share_rgsw(CRS, 4000 + j, j, total, ps[j][0], ps[j][2], ps[j][1], l, 1)
                  for j in range(total)]
        zero_l = np.stack([sh[0][0] for sh in shares])
        lead = o.ni_leader(l, total)
        for j in range(total):
            out = o.ni_produce(CRS, total, l, ksks[j], zero_l, shares[j][1][0], j == lead)
            e_exp = o.embedding * int(ps[j][1][l])
            if j == lead:
                check_rgsw(o, out, s, e_exp, (p.rlrg_da, p.rlrg_db), p.rlrg_logb, 1 << 13)
            else:
                check_rgsw(o, o.ntt_rows("backward", out), s, e_exp, (p.rgrg_da, p.rgrg_db),
                           p.rgrg_logb, 1 << 13)


@pytest.mark.parametrize("name,total", [("NI_2P", 2), ("NI_4P_HB_FR", 3)])
def test_aggregate_is_the_product_chain_and_decrypts(name, total):
    o = oracle_for(name)
    p = o.p
    ps = parties(o, total)
    s = sum(x[0].astype(np.int64) for x in ps)
    z = sum(x[1].astype(np.int64) for x in ps)
    ksks = uitos_ksks(o, total, ps)
    per = -(-p.lwe_n // total)
    l0, count = per - 1, 2  # straddles a leader boundary
    shares = [o.ni_share_rgsw(CRS, 4000 + j, j, total, ps[j][0], ps[j][2], ps[j][1], l0, count)
              for j in range(total)]
    zero = np.stack([sh[0] for sh in shares])
    cts = np.stack([sh[1] for sh in shares])
    got = o.ni_aggregate_rgsw(CRS, total, np.stack(ksks), l0, count, zero, cts)
    for t in range(count):
        l = l0 + t
        lead = o.ni_leader(l, total)
        want = o.ni_produce(CRS, total, l, ksks[lead], zero[:, t], cts[lead, t], True)
        for j in range(total):
            if j != lead:
                want = o.rgsw_by_rgsw(want, o.ni_produce(CRS, total, l, ksks[j], zero[:, t],
                                                         cts[j, t], False))
        assert np.array_equal(got[t], want)
        check_rgsw(o, got[t], s, o.embedding * int(z[l]), (p.rlrg_da, p.rlrg_db), p.rlrg_logb,
                   1 << 30)
