"""GPU: the non-interactive server key from party shares alone, through the
seeded wire formats (src/bool/keys.rs:1029-1063, :1240-1394), against the
oracle.

- Each party's share blob holds exactly the parts the lower-level share
  generators make for the same seeds (RGSW rows packed: the leader's rows
  with the RLWE x RGSW count, the other indices with the RGSW x RGSW count),
  and its u->s and automorphism-key / LWE-KSK parts equal the oracle's.
- The seeded server key aggregated from the shares (given in reverse party
  order) holds the u->s b parts (own b plus the others' zero encryptions),
  the aggregated RGSWs (sampled indices against the oracle) and the summed
  automorphism-key / LWE-KSK b parts.
- The server key expanded from the seeded key equals the oracle's expansion
  word for word, the parties' Ksk(u_j -> s) equal the oracle's, and gates on
  it decrypt correctly under the ideal secret.
- Shares of different seeds, two shares of one party and a share of another
  parameter set are refused.
"""
import sys

import numpy as np
import pytest

from conftest import oracle_for, rng

pytestmark = pytest.mark.gpu

CRS = bytes(range(90, 122))
PLAIN = {0: lambda a, b: 1 - (a & b), 4: lambda a, b: a ^ b}


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_ni_server_key_from_share_blobs_ni2p():
    import pzlib
    name, total = "NI_2P", 2
    o = oracle_for(name)
    p, N, n = o.p, o.N, o.p.lwe_n
    ctx = pzlib.Context(name)
    try:
        ps = []
        for j in range(total):
            s_j, z_j = ctx.keygen_secrets(1100 + j, 1)
            ps.append((s_j, z_j, ctx.keygen_user_secret(2100 + j)))
        log("ni wire: shares NI_2P")
        blobs = [ctx.ni_key_share(CRS, 3100 + j, j, total, ps[j][0], ps[j][2], ps[j][1])
                 for j in range(total)]
        per = -(-n // total)
        d1, dm = max(p.rlrg_da, p.rlrg_db), max(p.rgrg_da, p.rgrg_db)
        zeros, cts = [], []
        for j in range(total):
            h, sec = pzlib.ni_blob_sections(blobs[j])
            assert h.magic == b"PZNISHR" and h.user == j and h.total == total
            assert bytes(h.crs) == CRS and h.nsections == 7
            s_j, z_j, u_j = ps[j]
            kb, zo = ctx.ni_share_uitos(CRS, 3100 + j, j, total, s_j, u_j)
            wkb, wzo = o.ni_share_uitos(CRS, 3100 + j, j, total, s_j, u_j)
            assert np.array_equal(kb, wkb) and np.array_equal(zo, wzo)
            assert np.array_equal(sec[3], kb.reshape(-1)) and np.array_equal(sec[4], zo.reshape(-1))
            zr, ct = ctx.ni_share_rgsw(CRS, 3100 + j, j, total, s_j, u_j, z_j, 0, n)
            lead = np.arange(n)[(np.arange(n) // per) == j]
            rest = np.arange(n)[(np.arange(n) // per) != j]
            assert np.array_equal(sec[0], ct[lead][:, dm - d1:].reshape(-1))
            assert np.array_equal(sec[1], ct[rest].reshape(-1))
            assert np.array_equal(sec[2], zr.reshape(-1))
            assert not ct[lead][:, :dm - d1].any()
            ab, kk = o.ni_share_auto_ksk(CRS, 3100 + j, s_j, z_j)
            assert np.array_equal(sec[5], ab.reshape(-1)) and np.array_equal(sec[6], kk)
            zeros.append(zr)
            cts.append(ct)
        log("ni wire: aggregate")
        seeded = ctx.ni_aggregate_shares(blobs[::-1])
        h, sec = pzlib.ni_blob_sections(seeded)
        assert h.magic == b"PZNIKEY" and h.total == total and bytes(h.crs) == CRS
        secs = [pzlib.ni_blob_sections(b)[1] for b in blobs]
        uw = p.uitos_d * N
        oks = []
        for j in range(total):
            b = secs[j][3].astype(object)
            for k in range(total):
                if k != j:
                    slot = j if j < k else j - 1
                    b = b + secs[k][4][slot * uw:(slot + 1) * uw].astype(object)
            b = (b % o.Q).astype(np.uint64)
            assert np.array_equal(sec[0][j * uw:(j + 1) * uw], b), j
            zf = np.stack([secs[k][4][(j if j < k else j - 1) * uw:((j if j < k else j - 1) + 1)
                                      * uw] for k in range(total) if k != j])
            oks.append(o.ni_uitos_ksk(CRS, total, j, secs[j][3], zf))
        rows = 2 * (p.rlrg_da + p.rlrg_db)
        rgsw = sec[1].reshape(n, rows, N)
        zs, cs = np.stack(zeros), np.stack(cts)
        for l in (0, per - 1, per, n - 1):
            want = o.ni_aggregate_rgsw(CRS, total, np.stack(oks), l, 1, zs[:, l:l + 1],
                                       cs[:, l:l + 1])
            assert np.array_equal(rgsw[l], want[0]), l
        wa, wk = o.ni_aggregate_auto_ksk(np.stack([x[5].reshape(p.w + 1, p.auto_d, N)
                                                   for x in secs]),
                                         np.stack([x[6] for x in secs]))
        assert np.array_equal(sec[2], wa.reshape(-1)) and np.array_equal(sec[3], wk)
        log("ni wire: server key from the seeded key")
        key, ksks = ctx.ni_server_key_from_seeded(seeded, total)
        okey = o.ni_expand_server_key(CRS, rgsw, wa, wk)
        g_rgsw, g_auto, g_ksk = ctx.key_download(key)
        assert np.array_equal(g_rgsw.reshape(-1), okey.rgsw)
        assert np.array_equal(g_auto.reshape(-1), okey.auto_key)
        assert np.array_equal(g_ksk.reshape(-1), okey.lwe_ksk)
        for j in range(total):
            assert np.array_equal(ksks[j].download(), oks[j]), j
            ksks[j].free()
        s = sum(x[0].astype(np.int64) for x in ps).astype(np.int32)
        B = 128
        r = rng(88)
        b0 = r.integers(0, 2, size=B).astype(np.uint8)
        b1 = r.integers(0, 2, size=B).astype(np.uint8)
        ops = np.where(r.integers(0, 2, size=B) == 0, 0, 4).astype(np.uint32)
        c0, c1 = ctx.encrypt_bools(s, b0, 7200), ctx.encrypt_bools(s, b1, 9200)
        dout = ctx.array((B, N + 1))
        log("ni wire: gates on the key")
        ctx.gates(key, ctx.to_device(ops), ctx.to_device(c0), ctx.to_device(c1), dout, B)
        got = dout.download()
        want = np.array([PLAIN[int(ops[i])](int(b0[i]), int(b1[i])) for i in range(B)],
                        dtype=np.uint8)
        assert np.array_equal(ctx.decrypt_bools(s, got), want)
        for i in (0, B - 1):
            assert np.array_equal(got[i], o.gate(okey, int(ops[i]), c0[i], c1[i])), i
        key.free()
        # refusals
        other = ctx.ni_key_share(bytes(32), 3100, 1, total, ps[1][0], ps[1][2], ps[1][1])
        with pytest.raises(pzlib.PzError, match="seeds"):
            ctx.ni_aggregate_shares([blobs[0], other])
        with pytest.raises(pzlib.PzError, match="same party"):
            ctx.ni_aggregate_shares([blobs[0], blobs[0]])
        with pytest.raises(pzlib.PzError):
            ctx.ni_server_key_from_seeded(blobs[0])  # a share is not a seeded key
    finally:
        ctx.close()
    hb = pzlib.Context("NI_4P_HB_FR")
    try:
        s4, z4 = hb.keygen_secrets(1300, 1)
        u4 = hb.keygen_user_secret(1301)
        foreign = hb.ni_key_share(CRS, 1302, 1, total, s4, u4, z4)
    finally:
        hb.close()
    ctx = pzlib.Context(name)
    try:
        with pytest.raises(pzlib.PzError, match="parameter set"):
            ctx.ni_aggregate_shares([blobs[0], foreign])
    finally:
        ctx.close()
