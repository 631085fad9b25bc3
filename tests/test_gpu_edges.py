"""Edge inputs and error paths of the C ABI on the GPU.

- Empty batches (B = 0) of every batched entry point return success and leave
  the output untouched.
- Null buffers and out-of-range arguments return PZ_ERR_ARG, parameter sets
  without the needed decomposer PZ_ERR_UNSUPPORTED; nothing is launched.
- Keys and Ksk(u -> s) handles serve every context of their parameter set on
  their device (outputs equal the owner's), and are refused by a context of
  another parameter set (the kernels would index past the key).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import keys_for

pytestmark = pytest.mark.gpu

PZ_ERR_ARG, PZ_ERR_UNSUPPORTED = -1, -3


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import pzlib
    ctx = pzlib.Context("NI_2P")
    o, s, z, key = keys_for("NI_2P")
    k = ctx.key_upload(key.rgsw, key.auto_key, key.lwe_ksk)
    yield pzlib, ctx, k, s
    k.free()
    ctx.close()


def test_empty_batches(env):
    pzlib, ctx, k, s = env
    log("edges: empty batches")
    N, n = ctx.N, ctx.n
    sentinel = np.full((2, N + 1), 7, dtype=np.uint64)
    d = ctx.to_device(sentinel)
    d_u32 = ctx.to_device(np.zeros(2, dtype=np.uint32))
    d_u8 = ctx.to_device(np.zeros(2, dtype=np.uint8))
    L, h = ctx.L, ctx.h
    assert L.pz_gates(h, k.h, d_u32.ptr, d.ptr, d.ptr, d.ptr, 0, None) == 0
    assert L.pz_pbs(h, k.h, d_u32.ptr, d.ptr, 0, None) == 0
    assert L.pz_not(h, d.ptr, 0, None) == 0
    assert L.pz_lwe_key_switch(h, k.h, d.ptr, d.ptr, 0, None) == 0
    assert L.pz_rlwe_by_rgsw(h, k.h, d.ptr, d_u32.ptr, d_u8.ptr, 0, None) == 0
    assert L.pz_rlwe_auto(h, k.h, d.ptr, d_u32.ptr, d_u8.ptr, 0, None) == 0
    assert L.pz_ntt_forward(h, d.ptr, 0, None) == 0
    assert L.pz_sample_extract(h, d.ptr, d_u32.ptr, 0, d.ptr, None) == 0
    assert L.pz_rgsw_by_rgsw(h, d.ptr, d.ptr, 0, None) == 0
    ctx.sync()
    assert np.array_equal(d.download(), sentinel)


def test_out_of_range_key_indices_leave_rlwes_unchanged(env):
    """RLWE x RGSW / automorphism entries whose key index does not exist are
    skipped (no key read past the end); the other entries of the batch are
    the oracle's words."""
    pzlib, ctx, k, s = env
    log("edges: out-of-range key indices")
    o, _, _, key = keys_for("NI_2P")
    N, n, w = ctx.N, ctx.n, o.p.w
    r = np.random.default_rng(17)
    x = r.integers(0, ctx.Q, size=(4, 2, N), dtype=np.uint64)
    triv = np.zeros(4, dtype=np.uint8)
    for op, good, bad in (("ext", 5, n), ("auto", 3, w + 1)):
        idx = np.array([good, bad, bad + 1000, 0xFFFFFFFF], dtype=np.uint32)
        d = ctx.to_device(x)
        if op == "ext":
            ctx.rlwe_by_rgsw(k, d, ctx.to_device(idx), ctx.to_device(triv), 4)
            ct, cts = key.rgsw_ct(good)
            want0 = o.rlwe_by_rgsw_shoup(x[0], ct, cts, False)
        else:
            ctx.rlwe_auto(k, d, ctx.to_device(idx), ctx.to_device(triv), 4)
            kt, kts = key.auto_ct(good)
            want0 = o.rlwe_auto_shoup(x[0], kt, kts, good, False)
        got = d.download()
        assert np.array_equal(got[0], want0), op
        assert np.array_equal(got[1:], x[1:]), op


def test_null_and_range_errors(env):
    pzlib, ctx, k, s = env
    log("edges: null and range errors")
    L, h = ctx.L, ctx.h
    d = ctx.array((4, ctx.N + 1))
    d_u32 = ctx.to_device(np.zeros(4, dtype=np.uint32))
    assert L.pz_gates(h, k.h, None, d.ptr, d.ptr, d.ptr, 4, None) == PZ_ERR_ARG
    assert L.pz_gates(h, None, d_u32.ptr, d.ptr, d.ptr, d.ptr, 4, None) == PZ_ERR_ARG
    assert L.pz_pbs(h, k.h, d_u32.ptr, None, 4, None) == PZ_ERR_ARG
    assert L.pz_rgsw_by_rgsw(h, None, d.ptr, 1, None) == PZ_ERR_ARG
    crs = np.zeros(32, dtype=np.uint8)
    ks = (C.c_void_p * 2)(None, None)
    assert L.pz_ni_aggregate_rgsw(h, crs.ctypes.data, 2, ks, 0, 1, d.ptr, d.ptr, d.ptr,
                                  None) == PZ_ERR_ARG  # null Ksk handles
    assert L.pz_ni_aggregate_rgsw(h, crs.ctypes.data, 9, ks, 0, 1, d.ptr, d.ptr, d.ptr,
                                  None) == PZ_ERR_ARG  # more than 8 parties
    sec = np.zeros(ctx.N, dtype=np.int32)
    z = np.zeros(ctx.n, dtype=np.int32)
    out = np.zeros((1, 10, ctx.N), dtype=np.uint64)
    assert L.pz_ni_share_rgsw(h, crs.ctypes.data, crs.ctypes.data, 2, 2, sec.ctypes.data,
                              sec.ctypes.data, z.ctypes.data, 0, 1, out.ctypes.data,
                              out.ctypes.data) == PZ_ERR_ARG  # self >= total
    assert L.pz_ni_share_rgsw(h, crs.ctypes.data, crs.ctypes.data, 0, 2, sec.ctypes.data,
                              sec.ctypes.data, z.ctypes.data, ctx.n, 1, out.ctypes.data,
                              out.ctypes.data) == PZ_ERR_ARG  # index past lwe_n
    sp = pzlib.Context("SP_TEST")
    try:
        assert sp.L.pz_ni_share_uitos(sp.h, crs.ctypes.data, crs.ctypes.data, 0, 2,
                                      sec.ctypes.data, sec.ctypes.data, out.ctypes.data,
                                      out.ctypes.data) == PZ_ERR_UNSUPPORTED
    finally:
        sp.close()


def _gates(ctx, key, s, seed):
    r = np.random.default_rng(seed)
    B = 64
    b0 = r.integers(0, 2, B).astype(np.uint8)
    b1 = r.integers(0, 2, B).astype(np.uint8)
    c0, c1 = ctx.encrypt_bools(s, b0, seed), ctx.encrypt_bools(s, b1, seed + 1)
    out = ctx.array((B, ctx.N + 1))
    ctx.gates(key, ctx.to_device(np.zeros(B, np.uint32)), ctx.to_device(c0), ctx.to_device(c1),
              out, B)
    return out.download(), 1 - (b0 & b1)


def test_keys_across_contexts(env):
    pzlib, ctx, k, s = env
    log("edges: keys across contexts")
    first, want = _gates(ctx, k, s, 500)
    other = pzlib.Context("NI_2P")
    try:
        again, _ = _gates(other, k, s, 500)  # same parameter set: the key serves
        assert np.array_equal(again, first)
        assert np.array_equal(ctx.decrypt_bools(s, first), want)
    finally:
        other.close()
    wrong = pzlib.Context("NI_8P")
    try:
        B = 4
        d = wrong.array((B, wrong.N + 1))
        ops = wrong.to_device(np.zeros(B, np.uint32))
        with pytest.raises(pzlib.PzError):
            wrong.gates(k, ops, d, d, d, B)
        with pytest.raises(pzlib.PzError):
            wrong.pbs(k, ops, d, B)
        circ = pzlib.Circuit(wrong)
        circ.gate("nand", *circ.inputs(2))
        with pytest.raises(pzlib.PzError):
            circ.run(k, wrong.array((2, wrong.N + 1)))
        circ.close()
    finally:
        wrong.close()


def test_ksk_across_contexts():
    import pzlib
    log("edges: u->s key switching keys across contexts")
    a = pzlib.Context("NI_2P")
    b = pzlib.Context("NI_2P")
    c = pzlib.Context("NI_8P")
    try:
        s, z = a.keygen_secrets(900, 2)
        u = a.keygen_user_secret(901)
        ksk = a.keygen_rlwe_ksk(902, u, s)
        cts = a.encrypt_rlwe_bools(u, np.ones(a.N, dtype=np.uint8), 903)
        outs = []
        for ctx in (a, b):
            d_in, d_out = ctx.to_device(cts), ctx.array(cts.shape)
            ctx.rlwe_key_switch(ksk, d_in, d_out, 1)
            outs.append(d_out.download())
        assert np.array_equal(outs[0], outs[1])
        with pytest.raises(pzlib.PzError):
            c.rlwe_key_switch(ksk, c.to_device(cts), c.array(cts.shape), 1)
        ksk.free()
    finally:
        a.close()
        b.close()
        c.close()


def test_graph_cache_after_key_freed_elsewhere(env):
    """A key used (graph-captured) on a second context, then freed by its
    owner, then replaced by another key: the second context's next gates on
    the same buffers use the new key (the cache is keyed by key id, never by
    an address that a new key can reuse) and equal the oracle's words."""
    import pzlib
    pzlib_, ctx, k, s = env
    log("edges: graph cache after a key is freed on another context")
    o, s0, z0, key = keys_for("NI_2P")
    _, _, _, key2 = keys_for("NI_2P", seed_key=7)  # same secret, other key material
    other = pzlib.Context("NI_2P")
    try:
        r = np.random.default_rng(41)
        B = 4
        b0 = r.integers(0, 2, B).astype(np.uint8)
        b1 = r.integers(0, 2, B).astype(np.uint8)
        c0, c1 = other.encrypt_bools(s0, b0, 610), other.encrypt_bools(s0, b1, 620)
        d0, d1 = other.to_device(c0), other.to_device(c1)
        ops = np.array([0, 4, 0, 4], dtype=np.uint32)
        dops, dout = other.to_device(ops), other.array((B, other.N + 1))
        ka = ctx.key_upload(key.rgsw, key.auto_key, key.lwe_ksk)
        other.gates(ka, dops, d0, d1, dout, B)  # captures a graph with ka on `other`
        first = dout.download()
        ka.free()
        kb = ctx.key_upload(key2.rgsw, key2.auto_key, key2.lwe_ksk)
        other.gates(kb, dops, d0, d1, dout, B)
        got = dout.download()
        for i in range(B):
            assert np.array_equal(got[i], o.gate(key2, int(ops[i]), c0[i], c1[i])), i
            assert np.array_equal(first[i], o.gate(key, int(ops[i]), c0[i], c1[i])), i
        kb.free()
    finally:
        other.close()


def test_circuit_read_refuses_wires_added_after_run(env):
    """Reading a wire added after the last run is refused (PZ_ERR_STATE): the
    pool holds no value for it and may not hold its row; the wires of the run
    stay readable, and after another run the new wire is readable too."""
    import pzlib
    pzlib_, ctx, k, s = env
    log("edges: circuit read of a wire added after the run")
    circ = pzlib.Circuit(ctx)
    w = circ.inputs(2)
    g = circ.gate("nand", int(w[0]), int(w[1]))
    cts = ctx.encrypt_bools(s, np.array([1, 1], dtype=np.uint8), 630)
    d_in = ctx.to_device(cts)
    with pytest.raises(pzlib.PzError):
        circ.read([g])  # never run
    circ.run(k, d_in)
    assert ctx.decrypt_bools(s, circ.read([g]))[0] == 0
    extra = [circ.gate("xor", g, int(w[0])) for _ in range(80)]  # past the pool's 64 rows
    with pytest.raises(pzlib.PzError) as e:
        circ.read([extra[-1]])
    assert "after the last run" in str(e.value)
    assert ctx.decrypt_bools(s, circ.read([g]))[0] == 0
    circ.run(k, d_in)
    assert ctx.decrypt_bools(s, circ.read([extra[-1]]))[0] == 1  # NAND(1,1) xor 1
    circ.close()


def test_timed_gates_equal_untimed(env):
    """pz_gates with timing on replays a graph that carries HIP events around
    each kernel: the outputs equal the untimed replay's (and the oracle's),
    every kernel time is positive, and the three add up to less than the
    call's own wall time."""
    import time
    pzlib_, ctx, k, s = env
    log("edges: timed graph replay")
    o, _, _, key = keys_for("NI_2P")
    r = np.random.default_rng(43)
    B = 64
    b0 = r.integers(0, 2, B).astype(np.uint8)
    b1 = r.integers(0, 2, B).astype(np.uint8)
    c0, c1 = ctx.encrypt_bools(s, b0, 630), ctx.encrypt_bools(s, b1, 640)
    d0, d1 = ctx.to_device(c0), ctx.to_device(c1)
    ops = (np.arange(B) % 6).astype(np.uint32)
    dops, dout = ctx.to_device(ops), ctx.array((B, ctx.N + 1))
    ctx.gates(k, dops, d0, d1, dout, B)
    untimed = dout.download()
    ctx.set_timing(True)
    try:
        for rep in range(3):  # the first call captures the timed graph
            ctx.sync()
            t0 = time.perf_counter()
            ctx.gates(k, dops, d0, d1, dout, B)
            wall_ms = (time.perf_counter() - t0) * 1e3
            tim = ctx.last_timing()
            assert all(v > 0 for v in tim.values()), tim
            assert sum(tim.values()) <= wall_ms, (tim, wall_ms)
            assert np.array_equal(dout.download(), untimed), rep
    finally:
        ctx.set_timing(False)
    for i in range(0, B, 16):
        assert np.array_equal(untimed[i], o.gate(key, int(ops[i]), c0[i], c1[i])), i
