"""GPU non-interactive server-key shares and aggregation against the oracle,
word for word, and end to end.

- Party shares (u -> s parts and RGSW parts) equal the oracle's for the same
  seeds and secrets (NI_2P with 2 parties, NI_4P_HB_FR with 3, NI_4P_LB_SR
  with 4, NI_8P with 8).
- Ksk(u_j -> s) from the shares equals the oracle's.
- The aggregated RGSWs of indices straddling a leader boundary equal the
  oracle's (every produced RLWE and the product chain).
- Automorphism-key and LWE-KSK shares, their sums and the expansion of the
  seeded server key equal the oracle's (NI_2P with 2 parties, NI_8P with 3).
- Full NI_2P aggregation, 2 parties, all 520 LWE indices on the device: a
  sample of indices equals the oracle; the server key built from the parties'
  shares alone (RGSWs aggregated, automorphism keys and LWE KSK summed, a parts
  expanded from the CRS) equals the oracle's expansion word for word and
  bootstraps NAND / XOR gates that decrypt correctly under the ideal secret.
"""
import sys

import numpy as np
import pytest

from conftest import oracle_for, rng

pytestmark = pytest.mark.gpu

CRS = bytes(range(7, 39))


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def party_secrets(ctx, total):
    out = []
    for j in range(total):
        s, z = ctx.keygen_secrets(1000 + j, 1)
        u = ctx.keygen_user_secret(2000 + j)
        out.append((s, z, u))
    return out


def gpu_ksks(ctx, total, shares):
    ksks = []
    for j in range(total):
        zf = None
        if total > 1:
            zf = np.stack([shares[k][1][j if j < k else j - 1] for k in range(total) if k != j])
            zf = ctx.to_device(zf)
        ksks.append(ctx.ni_uitos_ksk(CRS, total, j, ctx.to_device(shares[j][0]), zf))
    return ksks


@pytest.mark.parametrize("name,total", [("NI_2P", 2), ("NI_4P_HB_FR", 3), ("NI_4P_LB_SR", 4),
                                        ("NI_8P", 8)])
def test_ni_shares_ksk_and_aggregate_match_oracle(name, total):
    import pzlib
    o = oracle_for(name)
    ctx = pzlib.Context(name)
    try:
        ps = party_secrets(ctx, total)
        log("ni shares %s parties=%d" % (name, total))
        shares = []
        for j in range(total):
            s, z, u = ps[j]
            g = ctx.ni_share_uitos(CRS, 3000 + j, j, total, s, u)
            w = o.ni_share_uitos(CRS, 3000 + j, j, total, s, u)
            assert np.array_equal(g[0], w[0]) and np.array_equal(g[1], w[1]), j
            shares.append(g)
        ksks = gpu_ksks(ctx, total, shares)
        oks = []
        for j in range(total):
            zf = np.stack([shares[k][1][j if j < k else j - 1] for k in range(total) if k != j])
            oks.append(o.ni_uitos_ksk(CRS, total, j, shares[j][0], zf))
            assert np.array_equal(ksks[j].download(), oks[j]), j
        per = -(-o.p.lwe_n // total)
        l0, count = per - 1, 2
        rg = []
        for j in range(total):
            s, z, u = ps[j]
            g = ctx.ni_share_rgsw(CRS, 4000 + j, j, total, s, u, z, l0, count)
            w = o.ni_share_rgsw(CRS, 4000 + j, j, total, s, u, z, l0, count)
            assert np.array_equal(g[0], w[0]) and np.array_equal(g[1], w[1]), j
            rg.append(g)
        zero = np.stack([x[0] for x in rg])
        cts = np.stack([x[1] for x in rg])
        log("ni aggregate %s l0=%d count=%d" % (name, l0, count))
        d_out = ctx.array((count, 2 * (o.p.rlrg_da + o.p.rlrg_db), o.N))
        ctx.ni_aggregate_rgsw(CRS, total, ksks, l0, count, ctx.to_device(zero),
                              ctx.to_device(cts), d_out)
        want = o.ni_aggregate_rgsw(CRS, total, np.stack(oks), l0, count, zero, cts)
        assert np.array_equal(d_out.download(), want)
        for k in ksks:
            k.free()
    finally:
        ctx.close()


PLAIN = {0: lambda a, b: 1 - (a & b), 4: lambda a, b: a ^ b}


def test_ni_full_aggregation_ni2p_bootstraps():
    import pzlib
    o = oracle_for("NI_2P")
    p = o.p
    total, n, N = 2, o.p.lwe_n, o.N
    ctx = pzlib.Context("NI_2P")
    try:
        ps = party_secrets(ctx, total)
        s = sum(x[0].astype(np.int64) for x in ps).astype(np.int32)
        z = sum(x[1].astype(np.int64) for x in ps).astype(np.int32)
        log("ni full shares NI_2P n=%d" % n)
        shares = [ctx.ni_share_uitos(CRS, 3000 + j, j, total, ps[j][0], ps[j][2])
                  for j in range(total)]
        ksks = gpu_ksks(ctx, total, shares)
        rg = [ctx.ni_share_rgsw(CRS, 4000 + j, j, total, ps[j][0], ps[j][2], ps[j][1], 0, n)
              for j in range(total)]
        zero = np.stack([x[0] for x in rg])
        cts = np.stack([x[1] for x in rg])
        rows = 2 * (p.rlrg_da + p.rlrg_db)
        d_out = ctx.array((n, rows, N))
        log("ni full aggregate NI_2P")
        ctx.ni_aggregate_rgsw(CRS, total, ksks, 0, n, ctx.to_device(zero), ctx.to_device(cts),
                              d_out)
        got = d_out.download()
        oks = [k.download() for k in ksks]
        for l in (0, n // 2 - 1, n // 2, n - 1):
            want = o.ni_aggregate_rgsw(CRS, total, np.stack(oks), l, 1, zero[:, l:l + 1],
                                       cts[:, l:l + 1])
            assert np.array_equal(got[l], want[0]), l
        # automorphism keys and LWE KSK from the parties' shares, summed, and
        # the whole server key expanded from the seeded parts on the device
        log("ni full auto/ksk shares NI_2P")
        ak = [ctx.ni_share_auto_ksk(CRS, 6000 + j, ps[j][0], ps[j][1]) for j in range(total)]
        d_as = ctx.to_device(np.stack([x[0] for x in ak]))
        d_ks = ctx.to_device(np.stack([x[1] for x in ak]))
        d_ab = ctx.array((p.w + 1, p.auto_d, N))
        d_kb = ctx.array((N * p.lwe_d,))
        ctx.ni_aggregate_auto_ksk(total, d_as, d_ks, d_ab, d_kb)
        key = ctx.ni_expand_server_key(CRS, d_out, d_ab, d_kb)
        assert np.array_equal(d_out.download(), got)  # inputs unchanged
        okey = o.ni_expand_server_key(CRS, got, d_ab.download(), d_kb.download())
        g_rgsw, g_auto, g_ksk = ctx.key_download(key)
        assert np.array_equal(g_rgsw.reshape(-1), okey.rgsw)
        assert np.array_equal(g_auto.reshape(-1), okey.auto_key)
        assert np.array_equal(g_ksk.reshape(-1), okey.lwe_ksk)
        B = 256
        r = rng(77)
        b0 = r.integers(0, 2, size=B).astype(np.uint8)
        b1 = r.integers(0, 2, size=B).astype(np.uint8)
        ops = np.where(r.integers(0, 2, size=B) == 0, 0, 4).astype(np.uint32)
        c0 = ctx.encrypt_bools(s, b0, 7100)
        c1 = ctx.encrypt_bools(s, b1, 9100)
        dout = ctx.array((B, N + 1))
        log("ni full aggregate gates B=%d" % B)
        ctx.gates(key, ctx.to_device(ops), ctx.to_device(c0), ctx.to_device(c1), dout, B)
        want_bits = np.array([PLAIN[int(ops[i])](int(b0[i]), int(b1[i])) for i in range(B)],
                             dtype=np.uint8)
        assert np.array_equal(ctx.decrypt_bools(s, dout.download()), want_bits)
        key.free()
        for k in ksks:
            k.free()
    finally:
        ctx.close()


@pytest.mark.parametrize("name,total", [("NI_2P", 2), ("NI_8P", 3)])
def test_ni_auto_ksk_shares_sums_and_expansion_match_oracle(name, total):
    import pzlib
    o = oracle_for(name)
    p, N = o.p, o.N
    ctx = pzlib.Context(name)
    try:
        ps = party_secrets(ctx, total)
        log("ni auto/ksk shares %s parties=%d" % (name, total))
        sh = []
        for j in range(total):
            g = ctx.ni_share_auto_ksk(CRS, 7000 + j, ps[j][0], ps[j][1])
            w = o.ni_share_auto_ksk(CRS, 7000 + j, ps[j][0], ps[j][1])
            assert np.array_equal(g[0], w[0]) and np.array_equal(g[1], w[1]), j
            sh.append(g)
        d_as = ctx.to_device(np.stack([x[0] for x in sh]))
        d_ks = ctx.to_device(np.stack([x[1] for x in sh]))
        d_ab, d_kb = ctx.array((p.w + 1, p.auto_d, N)), ctx.array((N * p.lwe_d,))
        ctx.ni_aggregate_auto_ksk(total, d_as, d_ks, d_ab, d_kb)
        wa, wk = o.ni_aggregate_auto_ksk(np.stack([x[0] for x in sh]), np.stack([x[1] for x in sh]))
        assert np.array_equal(d_ab.download(), wa) and np.array_equal(d_kb.download(), wk)
        rows = 2 * (p.rlrg_da + p.rlrg_db)
        rgsw = rng(31).integers(0, o.Q, size=(p.lwe_n, rows, N), dtype=np.uint64)
        log("ni expand server key %s" % name)
        key = ctx.ni_expand_server_key(CRS, ctx.to_device(rgsw), d_ab, d_kb)
        okey = o.ni_expand_server_key(CRS, rgsw, wa, wk)
        g_rgsw, g_auto, g_ksk = ctx.key_download(key)
        assert np.array_equal(g_rgsw.reshape(-1), okey.rgsw)
        assert np.array_equal(g_auto.reshape(-1), okey.auto_key)
        assert np.array_equal(g_ksk.reshape(-1), okey.lwe_ksk)
        key.free()
    finally:
        ctx.close()
