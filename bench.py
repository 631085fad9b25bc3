"""Bootstraps/sec of batched gate PBS on one MI355X (NonInteractiveLTE2Party).

A step is one batched launch of B independent NAND gate bootstraps (gate
linear step + f64 moddown + LWE key switch + persistent blind rotation +
sample extract) with keys and ciphertexts resident in HBM, replayed from a
captured hipGraph.  The plain baseline is the same path with one PBS per
launch.  The CPU baseline is the oracle (a C restatement of the reference's
Rust PBS) on the host cores.

Prints one JSON line (the driver's contract).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "phantom-zone_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

# MI355X: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz = 39.32e12 32-bit VALU lane-ops/s.
VALU_PEAK_OPS = 256 * 4 * 16 * 2.4e9
# VALU instructions of one 54-bit Shoup multiply (w*a - q*hi(w'*a)) in the
# gfx950 ISA of pz_arith.hpp's shoup_mul: 5 v_mad_u64_u32, 4 v_mul_lo_u32,
# 1 v_mul_hi_u32, 2 v_add3_u32, 1 v_lshl_add_u64, 2 v_sub(b), 4 v_mov.
VALU_OPS_PER_MODMUL = 19
HBM_PEAK_GBS = 8000.0


def modmul_per_pbs(N, n, autos):
    """Algorithmic 54-bit modmuls of one blind rotation (NI_2P figures:
    forward NTT = N/2*log2N, inverse = that + N, ext product = 2F + 2I + 4N,
    auto = F + 2I + 2N)."""
    logn = N.bit_length() - 1
    F = N // 2 * logn
    I = F + N
    ext = 2 * F + 2 * I + 4 * N
    aut = F + 2 * I + 2 * N
    return n * ext + autos * aut


def synthetic_key(p, seed):
    """Uniform key material of the right shapes (eval-domain RGSW / auto rows,
    KSK rows mod q_ks): the PBS does the same work for any key values."""
    r = np.random.default_rng(seed)
    N, n = p.rlwe_n, p.lwe_n
    rgsw = r.integers(0, p.rlwe_q, size=(n, 2 * p.rlrg_da + 2 * p.rlrg_db, N), dtype=np.uint64)
    auto = r.integers(0, p.rlwe_q, size=(p.w + 1, 2 * p.auto_d, N), dtype=np.uint64)
    ksk = r.integers(0, p.lwe_q, size=(N * p.lwe_d, n + 1), dtype=np.uint64)
    return rgsw, auto, ksk


def cpu_baseline(params, threads, per_thread):
    """Oracle PBS throughput on the host cores (one PBS stream per thread)."""
    from oracle import Oracle
    o = Oracle(params)
    s, z, key = o.keygen(0, 1, 2)
    count = threads * per_thread
    r = np.random.default_rng(5)
    lwes = r.integers(0, o.Q, size=(count, o.N + 1), dtype=np.uint64)
    _, secs = o.pbs_batch_threads(key, 0, lwes, threads)
    return {"value": count / secs, "unit": "bootstraps/sec", "cores": threads,
            "kind": "port",
            "sample": "%d NAND PBS at %s, %d threads x %d, %.1f s wall" %
                      (count, params, threads, per_thread, secs)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--params", default="NI_2P")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-per-thread", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box in this build has one MI355X; "
              "multi-GPU sharding is design-only (replicated keys, split gate batches).",
              file=sys.stderr)
        sys.exit(2)

    import pzlib
    ctx = pzlib.Context(args.params, device=0)
    p = ctx.p
    N, n, B = p.rlwe_n, p.lwe_n, args.batch
    rgsw, auto, ksk = synthetic_key(p, 7)
    key = ctx.key_upload(rgsw, auto, ksk)
    r = np.random.default_rng(3)
    c0 = r.integers(0, p.rlwe_q, size=(B, N + 1), dtype=np.uint64)
    c1 = r.integers(0, p.rlwe_q, size=(B, N + 1), dtype=np.uint64)
    ops = np.zeros(B, dtype=np.uint32)  # NAND
    d0, d1, dops = ctx.to_device(c0), ctx.to_device(c1), ctx.to_device(ops)
    dout = ctx.array((B, N + 1))

    for _ in range(args.warmup):
        ctx.gates(key, dops, d0, d1, dout, B)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.gates(key, dops, d0, d1, dout, B)
    ctx.sync()
    t1 = time.perf_counter()
    secs = t1 - t0
    value = B * args.steps / secs

    # per-kernel durations (HIP events on the launch stream) of one more step
    ctx.set_timing(True)
    ctx.gates(key, dops, d0, d1, dout, B)
    tim = ctx.last_timing()
    ctx.set_timing(False)
    autos = ctx.last_auto_counts(B)
    mm = float(sum(modmul_per_pbs(N, n, int(a)) for a in autos))
    br_s = tim["blind_rotate_ms"] / 1e3
    achieved = mm / br_s / 1e9  # Gmodmul/s
    peak = VALU_PEAK_OPS / VALU_OPS_PER_MODMUL / 1e9
    key_bytes = 8 * (rgsw.size + auto.size)
    ct_bytes = B * (4 * (n + 1) + 8 * (N + 1) + 4)
    roof_hbm = (key_bytes + ct_bytes) / br_s / 1e9

    # plain: one PBS per launch
    d0p, d1p, dopsp = ctx.to_device(c0[:1]), ctx.to_device(c1[:1]), ctx.to_device(ops[:1])
    doutp = ctx.array((1, N + 1))
    ctx.gates(key, dopsp, d0p, d1p, doutp, 1)
    ctx.sync()
    tp0 = time.perf_counter()
    for _ in range(args.plain_steps):
        ctx.gates(key, dopsp, d0p, d1p, doutp, 1)
        ctx.sync()
    plain = args.plain_steps / (time.perf_counter() - tp0)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(args.params, threads, args.cpu_per_thread)

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
        "data": "synthetic: uniform eval-domain keys and uniform input ciphertexts, seeded",
        "config": {"workload": "batched NAND PBS, %s (N=%d, n=%d), B=%d per launch" %
                               (args.params, N, n, B),
                   "batch": B, "params": args.params},
        "plain_bootstraps_per_sec": plain,
        "batched_over_plain": value / plain,
        "kernel_ms": tim,
        "autos_mean": float(np.mean(autos)),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak,
                     "unit": "Gmodmul/s", "frac": achieved / peak, "traffic": None,
                     "kernel": "k_blind_rotate",
                     "note": "54-bit Shoup modmuls/s of the blind rotation; peak = "
                             "VALU lane-op peak / %d ops per Shoup multiply" %
                             VALU_OPS_PER_MODMUL},
        "roofline_hbm": {"bound": "hbm", "achieved": roof_hbm, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": roof_hbm / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
