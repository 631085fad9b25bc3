"""GPU: the serving loop.  Requests from several sessions (different keys)
arrive over time and are batched across sessions level by level; every
result decrypts to the plaintext program result and equals, bit for bit, the
same request served alone.  A key budget smaller than the session count
forces evictions and re-admission."""
import numpy as np
import pytest

from conftest import oracle_for

pytestmark = pytest.mark.gpu


def make_sessions(ctx, name, count):
    out = []
    for i in range(count):
        s, z = ctx.keygen_secrets(100 + i, 1 if name == "SP_TEST" else 2)
        k = ctx.keygen_server(200 + i, s, z)
        rgsw, auto, ksk = ctx.key_download(k)
        k.free()
        out.append((s, z, rgsw, auto, ksk))
    return out


def submit_all(ctx, srv, sessions, progs, sess_ids, arrivals, seed):
    import pzlib
    rids, meta = [], []
    for j, (prog, si, t) in enumerate(zip(progs, sess_ids, arrivals)):
        circ = pzlib.Circuit(None)
        wires, nu, nb = prog.build(circ)
        cts = ctx.encrypt_bools(sessions[si][0], prog.input_bits(), seed + 1000 * j)
        rid = srv.submit(srv_sid(srv, si), circ, cts, wires, len(prog.steps), t)
        rids.append(rid)
        meta.append((nu, nb))
        circ.close()
    return rids, meta


def srv_sid(srv, si):
    return srv.sid_map[si]


def test_serving_mix_equals_alone():
    import pzlib
    import pzprog
    name = "SP_TEST"
    print("server mix %s" % name, flush=True)
    ctx = pzlib.Context(name)
    sessions = make_sessions(ctx, name, 3)
    r = np.random.default_rng(5)
    progs = [pzprog.Program(r, int(r.integers(1, 4))) for _ in range(9)]
    sess_ids = [j % 3 for j in range(9)]
    arrivals = [0.002 * j for j in range(9)]
    srv = pzlib.Server(ctx, hbm_budget_bytes=1 << 34, pool_rows=1 << 14, max_batch_pbs=4096)
    srv.sid_map = [srv.add_session(se[2], se[3], se[4]) for se in sessions]
    rids, meta = submit_all(ctx, srv, sessions, progs, sess_ids, arrivals, 40)
    srv.run()
    st = srv.stats()
    assert st["total_pbs"] == sum(srv.times(i)["pbs"] for i in rids)
    mixed = [srv.result(i) for i in rids]
    for j, prog in enumerate(progs):
        bits = ctx.decrypt_bools(sessions[sess_ids[j]][0], mixed[j])
        assert pzprog.decode(bits, *meta[j]) == prog.plain(), j
    # each request alone gives the same ciphertexts
    for j in (0, 4, 8):
        solo = pzlib.Server(ctx, hbm_budget_bytes=1 << 34, pool_rows=1 << 14)
        solo.sid_map = [solo.add_session(se[2], se[3], se[4]) for se in sessions]
        rj, _ = submit_all(ctx, solo, sessions, [progs[j]], [sess_ids[j]], [0.0], 40 + 1000 * j)
        solo.run()
        assert np.array_equal(solo.result(rj[0]), mixed[j])
        solo.close()
    srv.close()
    ctx.close()


def test_serving_budget_forces_eviction():
    import pzlib
    import pzprog
    name = "SP_TEST"
    print("server budget %s" % name, flush=True)
    ctx = pzlib.Context(name)
    sessions = make_sessions(ctx, name, 3)
    r = np.random.default_rng(6)
    progs = [pzprog.Program(r, 1, ops=("add", "lt")) for _ in range(6)]
    sess_ids = [0, 1, 2, 0, 1, 2]
    k = ctx.key_upload(sessions[0][2], sessions[0][3], sessions[0][4])
    one_key = k.nbytes() * 3 // 2  # room for a single key
    k.free()
    srv = pzlib.Server(ctx, hbm_budget_bytes=one_key, pool_rows=1 << 12)
    srv.sid_map = [srv.add_session(se[2], se[3], se[4]) for se in sessions]
    rids, meta = submit_all(ctx, srv, sessions, progs, sess_ids, [0.0] * 6, 70)
    srv.run()
    st = srv.stats()
    assert st["evictions"] >= 2 and st["key_uploads"] >= 3
    for j, prog in enumerate(progs):
        bits = ctx.decrypt_bools(sessions[sess_ids[j]][0], srv.result(rids[j]))
        assert pzprog.decode(bits, *meta[j]) == prog.plain(), j
    srv.close()
    ctx.close()
