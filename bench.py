"""Bootstraps/sec of batched gate PBS on one MI355X (NonInteractiveLTE2Party),
plus plain PBS, FheUint8 circuits and serving latency.

Headline (the JSON line's `value`): a step is one batched launch of B = 4096
independent NAND gate bootstraps (gate linear step + f64 moddown + LWE key
switch + persistent blind rotation + sample extract), keys and ciphertexts
resident in HBM, replayed from a captured hipGraph.  The key is the 2-party
NI_2P server key aggregated from the parties' seeded shares; inputs are
encryptions of seeded random bits under the ideal secret; outputs are
decrypted and checked after timing.

Extra fields: batch sweep (1024 / 16384), plain one-PBS-per-launch rate,
FheUint8 at NI_8P (256 mul + 256 div_rem + 256 lt, level by level), serving
(Poisson programs from 64 sessions; p50/p99 per-op latency; single-multiply
p99), the roofline of the blind-rotation kernel, and the CPU baseline (the
oracle -- a C restatement of the reference's Rust PBS -- on the host cores).
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "phantom-zone_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

# MI355X: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz = 39.32e12 32-bit VALU lane-ops/s.
VALU_PEAK_OPS = 256 * 4 * 16 * 2.4e9
# VALU instructions of one exact 54-bit Shoup multiply in the gfx950 ISA of
# pz_arith.hpp's two-step form (shoup_quot + shoup_rem) as first built:
# 6 v_mad_u64_u32, 3 v_mul_lo_u32, 1 v_mul_hi_u32, 2 v_lshl_add_u64,
# 1 v_add3_u32, 4 v_mov_b32.  The roofline keeps this divisor so fractions
# compare across builds; the exact form now takes 15 instructions (the cross
# products chained through v_mad_u64_u32 addends) and the [0,4q) form most
# of the blind rotation's products use 12.
VALU_OPS_PER_MODMUL = 17
HBM_PEAK_GBS = 8000.0
PLAIN = {0: lambda a, b: 1 - (a & b)}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def modmul_per_pbs(N, n, autos):
    """Algorithmic 54-bit modmuls of one blind rotation, counted as the
    reference computes it (src/ntt.rs, src/rgsw/runtime.rs): forward NTT
    F = N/2*log2(N), inverse I = F + N (its last stage scales by n^-1),
    RLWE x RGSW = 2F + 2I + 4N, automorphism = F + 2I + 2N.  (This build
    folds n^-1 into the key rows and skips those N per inverse; the work
    counted is the algorithm's, so the saving shows as achieved rate.)"""
    logn = N.bit_length() - 1
    F = N // 2 * logn
    I = F + N
    return n * (2 * F + 2 * I + 4 * N) + autos * (F + 2 * I + 2 * N)


NI_CRS = bytes(range(7, 39))


def keys(ctx, parties, seed=0):
    """The non-interactive server key of `parties` parties, from their seeded
    shares alone: each party draws its secrets (s_j, z_j, u_j) from seeds
    derived from `seed` and makes its share blob; the shares are aggregated
    into the seeded server key, which is expanded on the device.  Returns the
    ideal secrets s = sum s_j, z = sum z_j and the key."""
    blobs, s, z = [], None, None
    for j in range(parties):
        s_j, z_j = ctx.keygen_secrets(1000 * seed + 10 * j + 1, 1)
        u_j = ctx.keygen_user_secret(1000 * seed + 10 * j + 2)
        blobs.append(ctx.ni_key_share(NI_CRS, 1000 * seed + 10 * j + 3, j, parties, s_j, u_j, z_j))
        s = s_j.astype(np.int64) if s is None else s + s_j
        z = z_j.astype(np.int64) if z is None else z + z_j
    key, _ = ctx.ni_server_key_from_seeded(ctx.ni_aggregate_shares(blobs))
    return s.astype(np.int32), z.astype(np.int32), key


def batched(ctx, key, s, B, steps, warmup, seed, timed=False):
    """B NAND gates per hipGraph replay.  timed: the replays carry HIP events
    around each kernel (recorded inside the graph) and each replay is waited
    for, so the per-kernel times returned come from the timed replays
    themselves and add up to less than the step."""
    p = ctx.p
    r = np.random.default_rng(seed)
    b0 = r.integers(0, 2, size=B).astype(np.uint8)
    b1 = r.integers(0, 2, size=B).astype(np.uint8)
    c0 = ctx.encrypt_bools(s, b0, seed * 1000 + 1)
    c1 = ctx.encrypt_bools(s, b1, seed * 1000 + 1 + B)
    ops = np.zeros(B, dtype=np.uint32)  # NAND
    d0, d1, dops = ctx.to_device(c0), ctx.to_device(c1), ctx.to_device(ops)
    dout = ctx.array((B, p.rlwe_n + 1))
    ctx.set_timing(timed)
    for _ in range(warmup):
        ctx.gates(key, dops, d0, d1, dout, B)
    ctx.sync()
    tim = {"prep_ms": 0.0, "ks_ms": 0.0, "blind_rotate_ms": 0.0}
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.gates(key, dops, d0, d1, dout, B)
        if timed:
            for name, v in ctx.last_timing().items():
                tim[name] += v / steps
    ctx.sync()
    secs = time.perf_counter() - t0
    ctx.set_timing(False)
    got = ctx.decrypt_bools(s, dout.download())
    ok = bool(np.array_equal(got, (1 - (b0 & b1)).astype(np.uint8)))
    return B * steps / secs, secs, ok, tim


def host_path(ctx, key, s, B, steps, seed):
    """Batches whose ciphertexts start and end in host memory: per step both
    input arrays are copied to the device, the gates run and the outputs are
    copied back (pageable numpy buffers, through the library's stream-ordered
    copies), so the rate includes the PCIe transfers."""
    p = ctx.p
    r = np.random.default_rng(seed)
    b0 = r.integers(0, 2, size=B).astype(np.uint8)
    b1 = r.integers(0, 2, size=B).astype(np.uint8)
    c0 = ctx.encrypt_bools(s, b0, seed * 1000 + 1)
    c1 = ctx.encrypt_bools(s, b1, seed * 1000 + 1 + B)
    d0, d1 = ctx.array((B, p.rlwe_n + 1)), ctx.array((B, p.rlwe_n + 1))
    dops = ctx.to_device(np.zeros(B, dtype=np.uint32))
    dout = ctx.array((B, p.rlwe_n + 1))

    def step():
        d0.upload(c0)
        d1.upload(c1)
        ctx.gates(key, dops, d0, d1, dout, B)
        return dout.download()

    step()
    t0 = time.perf_counter()
    for _ in range(steps):
        got = step()
    secs = time.perf_counter() - t0
    ok = bool(np.array_equal(ctx.decrypt_bools(s, got), (1 - (b0 & b1)).astype(np.uint8)))
    return {"bootstraps_per_sec": B * steps / secs, "ms_per_step": secs / steps * 1e3,
            "copied_bytes_per_step": 3 * B * (p.rlwe_n + 1) * 8, "host_memory": "pageable",
            "outputs_correct": ok}


def fheuint8_circuits(args):
    """256 mul + 256 div_rem + 256 lt on uniform u8 at NI_8P, one circuit,
    every ready gate of a level in one launch."""
    import pzlib
    ctx = pzlib.Context("NI_8P")
    s, z, key = keys(ctx, 8, seed=80)
    r = np.random.default_rng(81)
    circ = pzlib.Circuit(ctx)
    specs, outs, in_bits = [], [], []
    for op, count in (("mul", args.u8_count), ("div_rem", args.u8_count), ("lt", args.u8_count)):
        for _ in range(count):
            a, b = int(r.integers(0, 256)), int(r.integers(0, 256))
            wa, wb = circ.inputs(8), circ.inputs(8)
            in_bits += [(a >> i) & 1 for i in range(8)] + [(b >> i) & 1 for i in range(8)]
            outs.append(circ.u8(op, wa, wb))
            specs.append((op, a, b))
    cts = ctx.encrypt_bools(s, np.array(in_bits, dtype=np.uint8), 82)
    d_in = ctx.to_device(cts)
    st = circ.stats()
    sizes = circ.level_sizes()
    ctx.sync()
    t0 = time.perf_counter()
    circ.run(key, d_in)
    ctx.sync()
    secs = time.perf_counter() - t0
    flat = np.concatenate(outs)
    bits = ctx.decrypt_bools(s, circ.read(flat))
    pos, ok = 0, True
    for (op, a, b), w in zip(specs, outs):
        v = bits[pos:pos + len(w)]
        pos += len(w)
        if op == "mul":
            ok &= sum(int(v[i]) << i for i in range(8)) == (a * b) & 255
        elif op == "div_rem":
            q = sum(int(v[i]) << i for i in range(8))
            rr = sum(int(v[8 + i]) << i for i in range(8))
            ok &= (q, rr) == ((a // b, a % b) if b else (255, a))
        else:
            ok &= int(v[0]) == int(a < b)
    return {"params": "NI_8P", "ops": {"mul": args.u8_count, "div_rem": args.u8_count,
                                       "lt": args.u8_count},
            "pbs": st["pbs"], "levels": st["depth"], "max_level_pbs": int(sizes.max()),
            "seconds": secs, "bootstraps_per_sec": st["pbs"] / secs,
            "ops_per_sec": 3 * args.u8_count / secs, "all_correct": bool(ok)}


def plain_rate(ctx, key, s, steps):
    """One NAND PBS per launch, synchronised after each: bootstraps/s, and
    whether the output decrypts to NAND(1, 1) = 0."""
    N = ctx.p.rlwe_n
    c0 = ctx.encrypt_bools(s, np.array([1], dtype=np.uint8), 9)
    c1 = ctx.encrypt_bools(s, np.array([1], dtype=np.uint8), 10)
    d0p, d1p = ctx.to_device(c0), ctx.to_device(c1)
    dopsp = ctx.to_device(np.zeros(1, dtype=np.uint32))
    doutp = ctx.array((1, N + 1))
    ctx.gates(key, dopsp, d0p, d1p, doutp, 1)
    ctx.sync()
    tp0 = time.perf_counter()
    for _ in range(steps):
        ctx.gates(key, dopsp, d0p, d1p, doutp, 1)
        ctx.sync()
    rate = steps / (time.perf_counter() - tp0)
    return rate, int(ctx.decrypt_bools(s, doutp.download())[0]) == 0


SERVING_SETS = (("NI_2P", 2), ("NI_4P_HB_FR", 4))  # 2-party and 4-party aggregated keys


def serving_leg(args):
    """The serving measurement alone; the serial baseline is the one-PBS rate
    measured here on an NI_2P key."""
    import pzlib
    ctx = pzlib.Context("NI_2P")
    s, z, key = keys(ctx, 2)
    rate, _ = plain_rate(ctx, key, s, args.plain_steps)
    key.free()
    ctx.close()
    return serving(args, rate)


def serving(args, plain_rate):
    """Poisson arrivals of FheUint8 programs (1..40 ops of add/sub/mul/div/
    rem/lt/eq) from 64 sessions, alternately with 2-party NI_2P and 4-party
    NI_4P_HB_FR keys (one server, one batch per parameter set per step), then
    single-multiply requests.  A client thread submits each request at its
    arrival time while the server loop runs (the online path)."""
    import pzlib
    import pzprog
    ctxs = {name: pzlib.Context(name) for name, _ in SERVING_SETS}
    blobs = []  # (parameter set, secret, key rows)
    tk = time.perf_counter()
    for i in range(args.distinct_keys):
        name, parties = SERVING_SETS[i % len(SERVING_SETS)]
        ctx = ctxs[name]
        s, z, k = keys(ctx, parties, seed=300 + 2 * i)
        blobs.append((name, s, ctx.key_download(k)))
        k.free()
    log("serving: %d keys generated in %.1f s" % (len(blobs), time.perf_counter() - tk))
    out = {}
    for mode in ("mixed", "mul"):
        srv = pzlib.Server(ctxs[SERVING_SETS[0][0]], hbm_budget_bytes=args.hbm_budget << 30,
                           pool_rows=1 << 17, max_batch_pbs=args.max_batch)
        sids = []
        for i in range(args.sessions):
            name, _, rows = blobs[i % len(blobs)]
            sids.append(srv.add_session(*rows, ctx=ctxs[name]))
        r = np.random.default_rng(400 if mode == "mixed" else 401)
        rate = args.rate if mode == "mixed" else args.mul_rate
        count = args.programs if mode == "mixed" else args.mul_requests
        gaps, jobs, progs, metas = [], [], [], []
        for j in range(count):
            gaps.append(float(r.exponential(1.0 / rate)))
            si = int(r.integers(0, args.sessions))
            if mode == "mixed":
                prog = pzprog.Program(r, int(r.integers(1, 41)))
            else:
                prog = pzprog.Program(r, 1, ops=("mul",))
            circ = pzlib.Circuit(None)
            wires, nu, nb = prog.build(circ)
            name, sec, _ = blobs[si % len(blobs)]
            cts = ctxs[name].encrypt_bools(sec, prog.input_bits(), 5000 + 97 * j)
            jobs.append((sids[si], circ, cts, wires, len(prog.steps)))
            progs.append((prog, si))
            metas.append((nu, nb))
        # online: a client thread submits each request at its Poisson arrival
        # time while the loop serves the others (arrival stamped at submission)
        rids, errs = [None] * count, []

        def client():
            try:
                for j, job in enumerate(jobs):
                    time.sleep(gaps[j])
                    rids[j] = srv.submit(*job, -1.0)
            except Exception as e:  # noqa: BLE001 -- reported below
                errs.append(e)
            finally:
                srv.intake(False)

        srv.intake(True)
        th = threading.Thread(target=client)
        th.start()
        srv.run()
        th.join()
        if errs:
            raise errs[0]
        for job in jobs:
            job[1].close()
        log("serving %s: %d requests served online" % (mode, count))
        st = srv.stats()
        lat, per_op, ok = [], [], True
        for rid, (prog, si), meta in zip(rids, progs, metas):
            tm = srv.times(rid)
            lat.append(tm["finished"] - tm["arrival"])
            per_op.append(lat[-1] / tm["ops"])
            name, sec, _ = blobs[si % len(blobs)]
            bits = ctxs[name].decrypt_bools(sec, srv.result(rid))
            ok &= pzprog.decode(bits, *meta) == prog.plain()
        sets = {}
        for i in range(args.sessions):
            sets[blobs[i % len(blobs)][0]] = sets.get(blobs[i % len(blobs)][0], 0) + 1
        res = {"requests": count, "arrival_rate_per_s": rate, "sessions": args.sessions,
               "submission": "online: a client thread submits at each Poisson arrival while "
                             "the loop serves; latency = completion - submission",
               "sessions_per_parameter_set": sets,
               "distinct_keys": args.distinct_keys, "total_pbs": st["total_pbs"],
               "launches": st["batches"], "wall_s": st["wall_s"],
               "bootstraps_per_sec": st["total_pbs"] / st["wall_s"],
               "vs_serial_plain": st["total_pbs"] / st["wall_s"] / plain_rate,
               "latency_p50_s": float(np.percentile(lat, 50)),
               "latency_p99_s": float(np.percentile(lat, 99)),
               "per_op_latency_p50_s": float(np.percentile(per_op, 50)),
               "per_op_latency_p99_s": float(np.percentile(per_op, 99)),
               "key_uploads": st["key_uploads"], "evictions": st["evictions"],
               "all_correct": bool(ok)}
        out[mode] = res
        srv.close()
    for c in ctxs.values():
        c.close()
    return out


def kernel_batch():
    """4096 forward + inverse NTTs of uniform rows and 4096 RLWE x RGSW
    products with a uniform evaluation-domain RGSW key at NI_2P."""
    import pzlib
    ctx = pzlib.Context("NI_2P")
    p = ctx.p
    N, n, B = p.rlwe_n, p.lwe_n, 4096
    r = np.random.default_rng(7)
    x = r.integers(0, ctx.Q, size=(B, N), dtype=np.uint64)
    d = ctx.to_device(x)
    reps = 20
    ctx.ntt("forward", d, B)
    ctx.ntt("backward", d, B)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.ntt("forward", d, B)
        ctx.ntt("backward", d, B)
    ctx.sync()
    pair_s = (time.perf_counter() - t0) / reps
    round_trip = bool(np.array_equal(d.download(), x))
    logn = N.bit_length() - 1
    F = N // 2 * logn
    rows = 2 * (p.rlrg_da + p.rlrg_db)
    rgsw = r.integers(0, ctx.Q, size=(n, rows, N), dtype=np.uint64)
    auto = np.zeros((p.w + 1, 2 * p.auto_d, N), dtype=np.uint64)
    ksk = np.zeros((N * p.lwe_d, n + 1), dtype=np.uint64)
    key = ctx.key_upload(rgsw, auto, ksk)
    rl = ctx.to_device(r.integers(0, ctx.Q, size=(B, 2, N), dtype=np.uint64))
    idx = ctx.to_device(r.integers(0, n, size=B, dtype=np.uint32))
    triv = ctx.to_device(np.zeros(B, dtype=np.uint8))
    ctx.rlwe_by_rgsw(key, rl, idx, triv, B)
    ctx.sync()
    reps_e = 5
    t0 = time.perf_counter()
    for _ in range(reps_e):
        ctx.rlwe_by_rgsw(key, rl, idx, triv, B)
    ctx.sync()
    ext_s = (time.perf_counter() - t0) / reps_e
    ext_mm = B * (2 * F + 2 * (F + N) + 4 * N)  # the algorithm's count (n^-1 scaling included)
    key.free()
    ctx.close()
    return {"params": "NI_2P", "rows": B,
            "ntt_fwd_plus_inv_ms": pair_s * 1e3,
            "ntt_gbs": 4 * B * N * 8 / pair_s / 1e9,
            "ntt_round_trip_exact": round_trip,
            "ext_products_ms": ext_s * 1e3,
            "ext_products_per_sec": B / ext_s,
            "ext_gmodmul_per_sec": ext_mm / ext_s / 1e9,
            "ext_frac_of_valu_roofline": ext_mm / ext_s / 1e9 /
                                          (VALU_PEAK_OPS / VALU_OPS_PER_MODMUL / 1e9),
            "ext_rlwe_gbs": 2 * B * 2 * N * 8 / ext_s / 1e9}


def input_path():
    """Input path at NI_2P: 64 RLWEs of encrypted bits (131072 bits) under a
    party's secret, switched to the ideal secret, every bit extracted."""
    import pzlib
    ctx = pzlib.Context("NI_2P")
    N = ctx.N
    s, z = ctx.keygen_secrets(500, 2)
    u = ctx.keygen_user_secret(501)
    ksk = ctx.keygen_rlwe_ksk(502, u, s)
    R = 64
    bits = np.random.default_rng(503).integers(0, 2, size=R * N).astype(np.uint8)
    cts = ctx.encrypt_rlwe_bools(u, bits, 504)
    d_in = ctx.to_device(cts)
    d_sw = ctx.array((R, 2, N))
    count = R * N
    src = np.stack([np.arange(count) // N, np.arange(count) % N], axis=1).astype(np.uint32)
    d_src = ctx.to_device(src)
    d_lwe = ctx.array((count, N + 1))
    ctx.rlwe_key_switch(ksk, d_in, d_sw, R)
    ctx.sync()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.rlwe_key_switch(ksk, d_in, d_sw, R)
    ctx.sync()
    ks_s = (time.perf_counter() - t0) / reps
    ctx.sample_extract(d_sw, d_src, count, d_lwe)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.sample_extract(d_sw, d_src, count, d_lwe)
    ctx.sync()
    ex_s = time.perf_counter() - t0
    check = 4096
    lw = ctx.array((check, N + 1))
    ctx.sample_extract(d_sw, ctx.to_device(src[:check]), check, lw)
    ok = bool(np.array_equal(ctx.decrypt_bools(s, lw.download()), bits[:check]))
    ksk.free()
    ctx.close()
    return {"params": "NI_2P", "rlwes": R, "bits": count,
            "key_switch_ms": ks_s * 1e3, "rlwes_per_sec": R / ks_s,
            "extract_ms": ex_s * 1e3, "extract_gbs": count * (N + 1) * 8 / ex_s / 1e9,
            "bits_per_sec": count / (ks_s + ex_s), "sample_correct": ok}


def ni_shares(ctx, total, l0, count):
    """Party secrets and NI server-key shares (client side) for indices
    [l0, l0 + count)."""
    ps, uit, rg, ak = [], [], [], []
    for j in range(total):
        s_j, z_j = ctx.keygen_secrets(1000 + j, 1)
        u_j = ctx.keygen_user_secret(2000 + j)
        ps.append((s_j, z_j, u_j))
        uit.append(ctx.ni_share_uitos(NI_CRS, 3000 + j, j, total, s_j, u_j))
        rg.append(ctx.ni_share_rgsw(NI_CRS, 4000 + j, j, total, s_j, u_j, z_j, l0, count))
        ak.append(ctx.ni_share_auto_ksk(NI_CRS, 5000 + j, s_j, z_j))
    return ps, uit, rg, ak


def ni_aggregation(params="NI_2P", total=2):
    """Server-key aggregation for `total` parties over all LWE indices: the
    u->s key switching keys from the shares, every RGSW (produced RLWEs and
    the leader x other RGSW products) and the automorphism-key / LWE-KSK share
    sums, shares resident in HBM; then the evaluation-domain key expanded from
    the aggregate and a batch of gates on it, decrypted with the ideal
    secret."""
    import pzlib
    ctx = pzlib.Context(params)
    p, N = ctx.p, ctx.N
    n = p.lwe_n
    t0 = time.perf_counter()
    ps, uit, rg, ak = ni_shares(ctx, total, 0, n)
    share_s = (time.perf_counter() - t0) / total
    d_kb = [ctx.to_device(uit[j][0]) for j in range(total)]
    d_zf = [ctx.to_device(np.stack([uit[k][1][j if j < k else j - 1]
                                    for k in range(total) if k != j])) for j in range(total)]
    d_zero = ctx.to_device(np.stack([x[0] for x in rg]))
    d_cts = ctx.to_device(np.stack([x[1] for x in rg]))
    rows = 2 * (p.rlrg_da + p.rlrg_db)
    d_out = ctx.array((n, rows, N))
    d_as = ctx.to_device(np.stack([x[0] for x in ak]))
    d_ks = ctx.to_device(np.stack([x[1] for x in ak]))
    d_ab = ctx.array(ak[0][0].shape)
    d_kk = ctx.array(ak[0][1].shape)

    def run():
        ksks = [ctx.ni_uitos_ksk(NI_CRS, total, j, d_kb[j], d_zf[j]) for j in range(total)]
        ctx.ni_aggregate_rgsw(NI_CRS, total, ksks, 0, n, d_zero, d_cts, d_out)
        ctx.ni_aggregate_auto_ksk(total, d_as, d_ks, d_ab, d_kk)
        ctx.sync()
        return ksks

    for k in run():
        k.free()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        for k in run():
            k.free()
    agg_s = (time.perf_counter() - t0) / reps
    ctx.sync()
    t0 = time.perf_counter()
    key = ctx.ni_expand_server_key(NI_CRS, d_out, d_ab, d_kk)
    expand_s = time.perf_counter() - t0
    s = sum(x[0].astype(np.int64) for x in ps).astype(np.int32)
    B = 256
    r = np.random.default_rng(600)
    b0 = r.integers(0, 2, size=B).astype(np.uint8)
    b1 = r.integers(0, 2, size=B).astype(np.uint8)
    ops = np.zeros(B, dtype=np.uint32)
    dout = ctx.array((B, N + 1))
    ctx.gates(key, ctx.to_device(ops), ctx.to_device(ctx.encrypt_bools(s, b0, 601)),
              ctx.to_device(ctx.encrypt_bools(s, b1, 602)), dout, B)
    ok = bool(np.array_equal(ctx.decrypt_bools(s, dout.download()), 1 - (b0 & b1)))
    key.free()
    ctx.close()
    produced = n * (p.rlrg_da + p.rlrg_db + (total - 1) * (p.rgrg_da + p.rgrg_db))
    return {"params": params, "parties": total, "lwe_indices": n,
            "aggregate_ms": agg_s * 1e3, "produced_rlwes": produced,
            "expand_ms": expand_s * 1e3, "share_gen_ms_per_party": share_s * 1e3,
            "gates_correct": ok}


def cpu_share():
    """Host cores this process may run on: its affinity mask, capped by the
    cgroup CPU quota when one is set."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(quota) // int(period)))
    except (OSError, ValueError):
        pass
    return n


def cpu_baseline(params, threads, per_thread):
    """Oracle PBS throughput on the host cores (one PBS stream per thread)."""
    from oracle import Oracle
    o = Oracle(params)
    s, z, key = o.keygen(0, 1, 2)
    count = threads * per_thread
    r = np.random.default_rng(5)
    lwes = r.integers(0, o.Q, size=(count, o.N + 1), dtype=np.uint64)
    _, secs = o.pbs_batch_threads(key, 0, lwes, threads)
    _, secs1 = o.pbs_batch_threads(key, 0, lwes[:4], 1)
    agg = None
    if o.p.uitos_d and o.p.rgrg_da and params == "NI_2P":
        # server-key aggregation, one thread, on a sample of 4 LWE indices
        # straddling the leaders' boundary, scaled to all indices
        total, l0, cnt = 2, o.p.lwe_n // 2 - 2, 4
        ps = [o.keygen_party(1000 + j) + (o.keygen_user_secret(2000 + j),) for j in range(total)]
        uit = [o.ni_share_uitos(NI_CRS, 3000 + j, j, total, ps[j][0], ps[j][2])
               for j in range(total)]
        rg = [o.ni_share_rgsw(NI_CRS, 4000 + j, j, total, ps[j][0], ps[j][2], ps[j][1], l0, cnt)
              for j in range(total)]
        t0 = time.perf_counter()
        ks = [o.ni_uitos_ksk(NI_CRS, total, j, uit[j][0],
                             np.stack([uit[k][1][j if j < k else j - 1]
                                       for k in range(total) if k != j])) for j in range(total)]
        o.ni_aggregate_rgsw(NI_CRS, total, np.stack(ks), l0, cnt, np.stack([x[0] for x in rg]),
                            np.stack([x[1] for x in rg]))
        per_index = (time.perf_counter() - t0) / cnt
        agg = {"aggregate_ms_est": per_index * o.p.lwe_n * 1e3, "cores": 1,
               "sample": "%d of %d LWE indices, 2 parties, scaled" % (cnt, o.p.lwe_n)}
    return {"value": count / secs, "unit": "bootstraps/sec", "cores": threads,
            "host_cores": os.cpu_count(), "cpu_share": cpu_share(),
            "build": "gcc -O3 (oracle/Makefile)",
            "kind": "port", "single_thread_pbs_ms": secs1 / 4 * 1e3,
            "ni_aggregation_ni2p": agg,
            "sample": "%d NAND PBS at %s (C restatement of the reference's Rust PBS), "
                      "%d threads x %d, %.1f s wall, %.1f s CPU" %
                      (count, params, threads, per_thread, secs, secs * threads)}


def copy_probe():
    import torch
    x = torch.ones(1 << 25, dtype=torch.int64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(2):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 8 << 25, "copy_probe_bytes_written": 8 << 25,
                      "dispatches": 2}))


def measured_traffic(params, B):
    """HBM bytes per k_blind_rotate launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy
    calibration), when it was taken on the same workload."""
    path = os.path.join(ROOT, "profiles",
                        "k_blind_rotate_%s_b%d_pmc.md" % (params.lower().replace("_", ""), B))
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_dispatch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--params", default="NI_2P")
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--sweep", default="1024,16384")
    ap.add_argument("--u8-count", type=int, default=256)
    ap.add_argument("--sessions", type=int, default=64)
    ap.add_argument("--distinct-keys", type=int, default=64)
    ap.add_argument("--programs", type=int, default=48)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--mul-requests", type=int, default=128)
    ap.add_argument("--mul-rate", type=float, default=40.0)
    ap.add_argument("--max-batch", type=int, default=8192)
    ap.add_argument("--hbm-budget", type=int, default=64, help="GiB of resident keys")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-per-thread", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch heuristic (profiling runs)")
    ap.add_argument("--no-extras", action="store_true", help="headline, plain and roofline only")
    ap.add_argument("--leg", choices=["kernel_batch", "input_path", "ni_aggregation",
                                      "ni_aggregation_ni8p", "serving", "fheuint8"],
                    help="run one secondary measurement alone (for profiling)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB read + 256 MiB write elementwise kernel twice "
                         "(calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.leg:
        legs = {"kernel_batch": kernel_batch, "input_path": input_path,
                "ni_aggregation": ni_aggregation,
                "ni_aggregation_ni8p": lambda: ni_aggregation("NI_8P", 8),
                "serving": lambda: serving_leg(args),
                "fheuint8": lambda: fheuint8_circuits(args)}
        print(json.dumps({args.leg: legs[args.leg]()}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only; multi-GPU runs are not implemented "
              "(the design shards gate batches over GPUs with replicated keys).",
              file=sys.stderr)
        sys.exit(2)

    import pzlib
    ctx = pzlib.Context(args.params, device=0)
    p = ctx.p
    N, n, B = p.rlwe_n, p.lwe_n, args.batch
    s, z, key = keys(ctx, 2, seed=0)
    log("keys ready")
    # launch shapes chosen by the library's autotuner for each batch size used
    # (profiling runs skip it so that every blind-rotation dispatch is the batch)
    tuned = None
    if not args.no_autotune:
        tuned = {str(b): ctx.autotune(key, b) for b in
                 [1, B] + [int(x) for x in args.sweep.split(",") if x and not args.no_extras]}
        log("autotune: %s" % tuned)
    # per-kernel durations come from the same timed replays (HIP events
    # recorded inside the graph around each kernel)
    value, secs, ok, tim = batched(ctx, key, s, B, args.steps, args.warmup, seed=3, timed=True)
    log("batched B=%d: %.0f bootstraps/s" % (B, value))
    autos = ctx.last_auto_counts(B)
    mm = float(sum(modmul_per_pbs(N, n, int(a)) for a in autos))
    br_s = tim["blind_rotate_ms"] / 1e3
    achieved = mm / br_s / 1e9  # Gmodmul/s
    peak = VALU_PEAK_OPS / VALU_OPS_PER_MODMUL / 1e9
    key_bytes = key.nbytes()
    ct_bytes = B * (4 * (n + 1) + 8 * (N + 1) + 4)
    roof_hbm = (key_bytes + ct_bytes) / br_s / 1e9

    # plain: one PBS per launch
    plain, ok_plain = plain_rate(ctx, key, s, args.plain_steps)
    ok &= ok_plain
    log("plain: %.1f bootstraps/s" % plain)

    # measured VALU rates: the instruction forms of the Shoup multiply, and
    # the Shoup multiply itself over independent chains (its achievable peak)
    probe = {name: ctx.valu_probe(kind) for kind, name in
             enumerate(("add_u32", "mul_lo_u32", "mul_hi_u32", "mad_u64_u32"))}
    probe["lshl_add_u64"] = ctx.valu_probe(5)
    probe["shoup_mul_per_s"] = ctx.valu_probe(4, 1024)
    log("valu probe: %s" % probe)

    extras = {}
    if not args.no_extras:
        sweep = {}
        for bs in [int(x) for x in args.sweep.split(",") if x]:
            v, _, okb, _ = batched(ctx, key, s, bs, 2 if bs <= 4096 else 1, 1, seed=bs)
            sweep[str(bs)] = {"bootstraps_per_sec": v, "outputs_correct": okb}
            log("batched B=%d: %.0f bootstraps/s" % (bs, v))
        sweep[str(B)] = {"bootstraps_per_sec": value, "outputs_correct": ok}
        extras["batch_sweep"] = sweep
        extras["host_path"] = host_path(ctx, key, s, B, 2, seed=5)
        log("host path (PCIe included): %s" % extras["host_path"])
        key.free()
        ctx.close()
        extras["kernel_batch"] = kernel_batch()
        log("kernel batch: %s" % extras["kernel_batch"])
        extras["input_path"] = input_path()
        log("input path: %s" % extras["input_path"])
        extras["ni_aggregation"] = ni_aggregation()
        log("ni aggregation: %s" % extras["ni_aggregation"])
        extras["ni_aggregation_ni8p"] = ni_aggregation("NI_8P", 8)
        log("ni aggregation NI_8P: %s" % extras["ni_aggregation_ni8p"])
        extras["fheuint8_ni8p"] = fheuint8_circuits(args)
        log("fheuint8: %s" % extras["fheuint8_ni8p"])
        extras["serving"] = serving(args, plain)
        log("serving: %s" % extras["serving"])

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or cpu_share()
        cpu = cpu_baseline(args.params, threads, args.cpu_per_thread)
        log("cpu: %s" % cpu)

    out = {
        "metric": "bootstraps/sec (batched vs plain PBS, 1 MI355X) + p99 FheUint8 op latency",
        "value": value,
        "unit": "bootstraps/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: 2-party NI_2P server key aggregated from the parties' seeded "
                "shares (seed 0), encryptions of seeded random bits under the ideal secret; "
                "outputs decrypted and checked",
        "config": {"workload": "batched NAND PBS, %s (N=%d, n=%d), B=%d per launch" %
                               (args.params, N, n, B), "batch": B, "params": args.params},
        "outputs_correct": bool(ok),
        "autotune": tuned,
        "plain_bootstraps_per_sec": plain,
        "batched_over_plain": value / plain,
        "kernel_ms": tim,
        "timing": {"ms_per_step": "wall clock over the timed hipGraph replays, each waited "
                                  "for",
                   "kernel_ms": "HIP events recorded inside the same replayed graphs around "
                                "each kernel, averaged over the %d timed steps" % args.steps,
                   "roofline_achieved": "kernel_ms.blind_rotate_ms"},
        "autos_mean": float(np.mean(autos)),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak,
                     "unit": "Gmodmul/s", "frac": achieved / peak,
                     "traffic": measured_traffic(args.params, B),
                     "traffic_source": "not measured in this run: rocprofv3 --pmc FETCH_SIZE / "
                                       "WRITE_SIZE of this workload, scaled by the copy "
                                       "calibration (profiles/k_blind_rotate_%s_b%d_pmc.md)" %
                                       (args.params.lower().replace("_", ""), B),
                     "kernel": "k_blind_rotate",
                     "note": "54-bit Shoup modmuls/s of the blind rotation; peak = "
                             "VALU lane-op peak / %d ops per Shoup multiply" %
                             VALU_OPS_PER_MODMUL},
        "valu_probe": {"lane_ops_per_s": {k: v for k, v in probe.items() if k != "shoup_mul_per_s"},
                       "shoup_mul_per_s": probe["shoup_mul_per_s"],
                       "blind_rotate_frac_of_probe": mm / br_s / probe["shoup_mul_per_s"]},
        "roofline_hbm": {"bound": "hbm", "achieved": roof_hbm, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": roof_hbm / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
