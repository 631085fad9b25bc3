"""GPU, full batch sizes.

- The kernel-level batch: 4096 forward + inverse NTTs of uniform rows, and
  4096 RLWE x RGSW products of uniform RLWEs with a uniform evaluation-domain
  RGSW key (logB = 17, d_a = d_b = 1), every output word against the oracle.
- Batched gate bootstraps at B = 256, 1024, 4096 and 16384 (NAND / XOR mix,
  NI_2P),
  with the automatic kernel choice (split kernel, both variants) and with
  one work group per bootstrap forced:
  every output decrypts to the gate of the input bits, a sample of outputs
  equals the oracle word for word, and replays of the captured graph repeat
  the first launch exactly.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import keys_for, oracle_for, rng, uniform_mod

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import pzlib
    c = pzlib.Context("NI_2P")
    yield c
    c.close()


@pytest.mark.parametrize("pair", [("forward", "backward"), ("forward_lazy", "backward_lazy")])
def test_ntt_batch_4096(ctx, pair):
    o = oracle_for("NI_2P")
    B = 4096
    log("ntt batch %s B=%d" % (pair[0], B))
    x = uniform_mod(rng(101), o.Q, (B, o.N))
    d = ctx.to_device(x)
    ctx.ntt(pair[0], d, B)
    f = d.download()
    assert np.array_equal(f, o.ntt_rows(pair[0], x))
    ctx.ntt(pair[1], d, B)
    back = d.download()
    assert np.array_equal(back, o.ntt_rows(pair[1], f))
    if pair[1] == "backward":
        assert np.array_equal(back, x)  # strict round trip


def test_ext_product_batch_4096_uniform_rgsw(ctx):
    from oracle.oracle import OracleKey
    o, s, z, key = keys_for("NI_2P")
    r = rng(102)
    uk = OracleKey(o.sizes, o.N)
    uk.rgsw[:] = uniform_mod(r, o.Q, uk.rgsw.shape)
    uk.auto_key[:] = key.auto_key
    uk.lwe_ksk[:] = key.lwe_ksk
    o.L.pzo_key_shoup(o.ev, C.byref(uk.cstruct()))
    k = ctx.key_upload(uk.rgsw, uk.auto_key, uk.lwe_ksk)
    B = 4096
    log("ext product batch B=%d" % B)
    x = uniform_mod(r, o.Q, (B, 2, o.N))
    idx = r.integers(0, o.n, size=B, dtype=np.uint32)
    triv = np.zeros(B, dtype=np.uint8)
    d = ctx.to_device(x)
    ctx.rlwe_by_rgsw(k, d, ctx.to_device(idx), ctx.to_device(triv), B)
    got = d.download()
    for b in range(B):
        ct, cts = uk.rgsw_ct(int(idx[b]))
        assert np.array_equal(got[b], o.rlwe_by_rgsw_shoup(x[b], ct, cts, False)), b
    k.free()


PLAIN = {0: lambda a, b: 1 - (a & b), 4: lambda a, b: a ^ b}


@pytest.mark.parametrize("B,mode", [(1024, 0), (4096, 0), (16384, 0), (1024, 1), (256, 0),
                                    (4096, 3), (256, 3)])
def test_gates_full_batch(ctx, B, mode):
    o, s, z, key = keys_for("NI_2P")
    k = ctx.key_upload(key.rgsw, key.auto_key, key.lwe_ksk)
    log("gates full batch B=%d mode=%d" % (B, mode))
    ctx.set_blind_rotate_mode(mode)
    r = rng(103 + B)
    b0 = r.integers(0, 2, size=B).astype(np.uint8)
    b1 = r.integers(0, 2, size=B).astype(np.uint8)
    ops = np.where(r.integers(0, 2, size=B) == 0, 0, 4).astype(np.uint32)
    c0 = ctx.encrypt_bools(s, b0, 7000 + B)
    c1 = ctx.encrypt_bools(s, b1, 9000 + B)
    d0, d1, dops = ctx.to_device(c0), ctx.to_device(c1), ctx.to_device(ops)
    dout = ctx.array((B, o.N + 1))
    ctx.gates(k, dops, d0, d1, dout, B)
    first = dout.download()
    want_bits = np.array([PLAIN[int(ops[i])](int(b0[i]), int(b1[i])) for i in range(B)],
                         dtype=np.uint8)
    assert np.array_equal(ctx.decrypt_bools(s, first), want_bits)
    for i in [0, 1, B // 2, B - 1]:
        assert np.array_equal(first[i], o.gate(key, int(ops[i]), c0[i], c1[i])), i
    for _ in range(2):  # graph replays
        ctx.gates(k, dops, d0, d1, dout, B)
        assert np.array_equal(dout.download(), first)
    ctx.set_blind_rotate_mode(0)
    k.free()


@pytest.mark.parametrize("B", [1, 300, 1024])
def test_autotune_keeps_outputs(ctx, B):
    """The autotuner changes only the launch shape: outputs before and after
    it are identical, and its choice is a variant the device can run."""
    o, s, z, key = keys_for("NI_2P")
    k = ctx.key_upload(key.rgsw, key.auto_key, key.lwe_ksk)
    log("autotune B=%d" % B)
    r = rng(104 + B)
    b0 = r.integers(0, 2, size=B).astype(np.uint8)
    b1 = r.integers(0, 2, size=B).astype(np.uint8)
    ops = np.where(r.integers(0, 2, size=B) == 0, 0, 4).astype(np.uint32)
    d0 = ctx.to_device(ctx.encrypt_bools(s, b0, 11000 + B))
    d1 = ctx.to_device(ctx.encrypt_bools(s, b1, 12000 + B))
    dops, dout = ctx.to_device(ops), ctx.array((B, o.N + 1))
    ctx.gates(k, dops, d0, d1, dout, B)
    before = dout.download()
    choice = ctx.autotune(k, B)
    assert choice["variant"] in ("one_half", "split_latency", "split", "split4_latency",
                                 "split4")
    assert choice["groups_per_cu"] >= 1 and choice["ms_per_batch"] > 0
    ctx.gates(k, dops, d0, d1, dout, B)
    assert np.array_equal(dout.download(), before)
    k.free()
