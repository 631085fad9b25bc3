"""GPU non-interactive server-key shares and aggregation against the oracle,
word for word, and end to end.

- Party shares (u -> s parts and RGSW parts) equal the oracle's for the same
  seeds and secrets (NI_2P with 2 parties, NI_4P_HB_FR with 3, NI_4P_LB_SR
  with 4, NI_8P with 8).
- Ksk(u_j -> s) from the shares equals the oracle's.
- The aggregated RGSWs of indices straddling a leader boundary equal the
  oracle's (every produced RLWE and the product chain).
- Full NI_2P aggregation, 2 parties, all 520 LWE indices on the device: a
  sample of indices equals the oracle, and a server key whose RGSW part is the
  aggregate (automorphism keys and LWE KSK generated under the same ideal
  secret, as the reference's share sums give) bootstraps NAND / XOR gates that
  decrypt correctly under the ideal secret.
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
        # server key: aggregated RGSWs in the evaluation domain
        ctx.ntt("forward", d_out, n * rows)
        rgsw_eval = d_out.download()
        base = ctx.keygen_server(5, s, z)
        _, auto, ksk = ctx.key_download(base)
        base.free()
        key = ctx.key_upload(rgsw_eval, auto, ksk)
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
