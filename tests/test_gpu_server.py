"""GPU: the serving loop.  Requests from several sessions (different keys)
arrive over time and are batched across sessions level by level; every
result decrypts to the plaintext program result and equals, bit for bit, the
same request served alone and the oracle's evaluation of the same gates.  A
key budget smaller than the session count forces evictions and
re-admission.  64 sessions (32 two-party NI_2P and 32 four-party NI_4P_HB_FR
keys, all resident) serve Poisson arrivals."""
import numpy as np
import pytest

from conftest import oracle_cone, oracle_for, oracle_key_from_rows

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
    # the oracle's evaluation of the same gates gives the same ciphertexts
    o = oracle_for(name)
    for j in (1, 5):
        se = sessions[sess_ids[j]]
        okey = oracle_key_from_rows(o, se[2], se[3], se[4])
        circ = pzlib.Circuit(None)
        wires, _, _ = progs[j].build(circ)
        cts = ctx.encrypt_bools(se[0], progs[j].input_bits(), 40 + 1000 * j)
        want = oracle_cone(o, okey, circ.nodes(), cts, wires)
        for i, w in enumerate(wires):
            assert np.array_equal(mixed[j][i], want[int(w)]), (j, int(w))
        circ.close()
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


def test_serving_two_parameter_sets():
    """Sessions of two parameter sets (2-party NI_2P and 4-party NI_4P_HB_FR
    keys) in one server: each step launches one batch per parameter set on
    its context's stream; every result decrypts correctly and equals the
    request served alone by a server of its own parameter set."""
    import pzlib
    import pzprog
    print("server two parameter sets", flush=True)
    ctxs = {"NI_2P": pzlib.Context("NI_2P"), "NI_4P_HB_FR": pzlib.Context("NI_4P_HB_FR")}
    parties = {"NI_2P": 2, "NI_4P_HB_FR": 4}
    sessions = []  # (set name, secret, rgsw, auto, ksk)
    for i in range(4):
        name = "NI_2P" if i % 2 == 0 else "NI_4P_HB_FR"
        c = ctxs[name]
        s, z = c.keygen_secrets(500 + i, parties[name])
        k = c.keygen_server(600 + i, s, z)
        rgsw, auto, ksk = c.key_download(k)
        k.free()
        sessions.append((name, s, rgsw, auto, ksk))
    r = np.random.default_rng(9)
    progs = [pzprog.Program(r, 1, ops=("add", "lt", "eq")) for _ in range(8)]
    sess_ids = [j % 4 for j in range(8)]
    arrivals = [0.001 * j for j in range(8)]

    def serve(server_ctx_name, only=None):
        srv = pzlib.Server(ctxs[server_ctx_name], hbm_budget_bytes=1 << 34, pool_rows=1 << 13)
        sid = {}
        for si, se in enumerate(sessions):
            if only is None or se[0] == only:
                sid[si] = srv.add_session(se[2], se[3], se[4], ctx=ctxs[se[0]])
        rids = {}
        for j, (prog, si, t) in enumerate(zip(progs, sess_ids, arrivals)):
            if si not in sid:
                continue
            circ = pzlib.Circuit(None)
            wires, nu, nb = prog.build(circ)
            cts = ctxs[sessions[si][0]].encrypt_bools(sessions[si][1], prog.input_bits(), 77 + j)
            rids[j] = (srv.submit(sid[si], circ, cts, wires, len(prog.steps),
                                  t if only is None else 0.0), nu, nb)
            circ.close()
        srv.run()
        out = {j: (srv.result(rid), nu, nb) for j, (rid, nu, nb) in rids.items()}
        st = srv.stats()
        srv.close()
        return out, st

    mixed, st = serve("NI_2P")
    assert len(mixed) == 8
    for j, (ct, nu, nb) in mixed.items():
        name, s = sessions[sess_ids[j]][0], sessions[sess_ids[j]][1]
        bits = ctxs[name].decrypt_bools(s, ct)
        assert pzprog.decode(bits, nu, nb) == progs[j].plain(), j
    for name in ("NI_2P", "NI_4P_HB_FR"):
        alone, _ = serve(name, only=name)
        for j, (ct, _, _) in alone.items():
            assert np.array_equal(ct, mixed[j][0]), (name, j)
    for c in ctxs.values():
        c.close()


def test_server_refuses_other_ring_size():
    """A session of a context with another ring size cannot share the wire
    pool: PZ_ERR_UNSUPPORTED, and the server keeps working."""
    import pzlib
    print("server refuses other ring size", flush=True)
    big, small = pzlib.Context("NI_2P"), pzlib.Context("SP_TEST")
    s, z = small.keygen_secrets(1, 1)
    k = small.keygen_server(2, s, z)
    rgsw, auto, ksk = small.key_download(k)
    k.free()
    srv = pzlib.Server(big, hbm_budget_bytes=1 << 34, pool_rows=1 << 10)
    with pytest.raises(pzlib.PzError):
        srv.add_session(rgsw, auto, ksk, ctx=small)
    assert srv.stats()["key_uploads"] == 0
    srv.close()
    small.close()
    big.close()


def test_serving_64_sessions_poisson():
    """64 sessions, each with its own key resident (32 two-party NI_2P and 32
    four-party NI_4P_HB_FR), Poisson arrivals of programs of 1-6 FheUint8
    operations.  Every result decrypts to the program's plaintext result;
    four requests (two per parameter set) served alone give the same words;
    one single-comparison request equals the oracle's gates word for word."""
    import pzlib
    import pzprog
    sets = (("NI_2P", 2), ("NI_4P_HB_FR", 4))
    ctxs = {name: pzlib.Context(name) for name, _ in sets}
    sessions = []  # (set name, secret, rgsw, auto, ksk)
    for i in range(64):
        name, parties = sets[i % 2]
        c = ctxs[name]
        s, z = c.keygen_secrets(800 + i, parties)
        k = c.keygen_server(900 + i, s, z)
        rgsw, auto, ksk = c.key_download(k)
        k.free()
        sessions.append((name, s, rgsw, auto, ksk))
    print("serving 64 sessions: keys ready", flush=True)
    r = np.random.default_rng(64)
    count = 64
    progs, sess_ids, arrivals, t = [], [], [], 0.0
    for j in range(count):
        t += float(r.exponential(1.0 / 40.0))
        arrivals.append(t)
        sess_ids.append(int(r.integers(0, 64)))
        progs.append(pzprog.Program(r, int(r.integers(1, 7))))
    # request 0: one less-than at session 0 (NI_2P), checked against the oracle
    progs[0] = pzprog.Program(np.random.default_rng(65), 1, ops=("lt",))
    sess_ids[0] = 0

    def serve(jobs):
        srv = pzlib.Server(ctxs["NI_2P"], hbm_budget_bytes=1 << 36, pool_rows=1 << 16,
                           max_batch_pbs=8192)
        sid = [srv.add_session(se[2], se[3], se[4], ctx=ctxs[se[0]]) for se in sessions]
        rids = {}
        for j, arrival in jobs:
            si = sess_ids[j]
            circ = pzlib.Circuit(None)
            wires, nu, nb = progs[j].build(circ)
            cts = ctxs[sessions[si][0]].encrypt_bools(sessions[si][1], progs[j].input_bits(),
                                                      6000 + 13 * j)
            rids[j] = (srv.submit(sid[si], circ, cts, wires, len(progs[j].steps), arrival),
                       nu, nb, wires, cts)
            circ.close()
        srv.run()
        out = {j: (srv.result(rid), nu, nb, wires, cts)
               for j, (rid, nu, nb, wires, cts) in rids.items()}
        st = srv.stats()
        srv.close()
        return out, st

    mixed, st = serve(list(enumerate(arrivals)))
    print("serving 64 sessions: %d requests, %d bootstraps, %d launches, %.2f s" %
          (count, st["total_pbs"], st["batches"], st["wall_s"]), flush=True)
    assert st["key_uploads"] == 64 and st["evictions"] == 0
    for j, (ct, nu, nb, _, _) in mixed.items():
        name, s = sessions[sess_ids[j]][0], sessions[sess_ids[j]][1]
        assert pzprog.decode(ctxs[name].decrypt_bools(s, ct), nu, nb) == progs[j].plain(), j
    picks = [j for j in range(1, count) if sessions[sess_ids[j]][0] == "NI_2P"][:2] + \
            [j for j in range(1, count) if sessions[sess_ids[j]][0] == "NI_4P_HB_FR"][:2]
    alone, _ = serve([(j, 0.0) for j in picks])
    for j in picks:
        assert np.array_equal(alone[j][0], mixed[j][0]), j
    o = oracle_for("NI_2P")
    se = sessions[0]
    okey = oracle_key_from_rows(o, se[2], se[3], se[4])
    circ = pzlib.Circuit(None)
    wires, _, _ = progs[0].build(circ)
    print("serving 64 sessions: oracle cone of request 0", flush=True)
    want = oracle_cone(o, okey, circ.nodes(), mixed[0][4], wires)
    for i, w in enumerate(wires):
        assert np.array_equal(mixed[0][0][i], want[int(w)]), int(w)
    circ.close()
    for c in ctxs.values():
        c.close()


def test_online_submissions_from_another_thread():
    """Requests submitted by a second thread while the loop serves others
    (intake open, arrival stamped at submission): later requests arrive while
    earlier ones are on the GPU, every result decrypts correctly and equals
    the same request served alone."""
    import threading
    import time

    import pzlib
    import pzprog
    name = "NI_2P"
    print("server online submissions %s" % name, flush=True)
    ctx = pzlib.Context(name)
    sessions = make_sessions(ctx, name, 3)
    r = np.random.default_rng(12)
    count = 12
    progs = [pzprog.Program(r, int(r.integers(1, 3)), ops=("add", "lt", "mul")) for _ in range(count)]
    sess_ids = [j % 3 for j in range(count)]
    gaps = [float(g) for g in r.exponential(1.0 / 40.0, size=count)]
    built = []
    for j, prog in enumerate(progs):
        circ = pzlib.Circuit(None)
        wires, nu, nb = prog.build(circ)
        cts = ctx.encrypt_bools(sessions[sess_ids[j]][0], prog.input_bits(), 900 + 7 * j)
        built.append((circ, wires, nu, nb, cts))
    srv = pzlib.Server(ctx, hbm_budget_bytes=1 << 34, pool_rows=1 << 15)
    srv.sid_map = [srv.add_session(se[2], se[3], se[4]) for se in sessions]
    rids = [None] * count
    errors = []

    def producer():
        try:
            for j in range(count):
                time.sleep(gaps[j])
                circ, wires, _, _, cts = built[j]
                rids[j] = srv.submit(srv.sid_map[sess_ids[j]], circ, cts, wires,
                                     len(progs[j].steps), -1.0)
        except Exception as e:  # noqa: BLE001 -- reported by the main thread
            errors.append(e)
        finally:
            srv.intake(False)

    srv.intake(True)
    th = threading.Thread(target=producer)
    th.start()
    srv.run()  # returns once the intake is closed and every request is done
    th.join()
    assert not errors, errors
    times = [srv.times(i) for i in rids]
    # online: a request arrived while an earlier one was still being served
    assert any(times[j]["arrival"] < times[i]["finished"] for i in range(count)
               for j in range(i + 1, count))
    mixed = [srv.result(i) for i in rids]
    for j, prog in enumerate(progs):
        bits = ctx.decrypt_bools(sessions[sess_ids[j]][0], mixed[j])
        assert pzprog.decode(bits, built[j][2], built[j][3]) == prog.plain(), j
    for j in (0, count // 2, count - 1):
        solo = pzlib.Server(ctx, hbm_budget_bytes=1 << 34, pool_rows=1 << 14)
        sid = [solo.add_session(se[2], se[3], se[4]) for se in sessions]
        circ, wires, _, _, cts = built[j]
        rid = solo.submit(sid[sess_ids[j]], circ, cts, wires, len(progs[j].steps), 0.0)
        solo.run()
        assert np.array_equal(solo.result(rid), mixed[j]), j
        solo.close()
    for b in built:
        b[0].close()
    srv.close()
    ctx.close()
