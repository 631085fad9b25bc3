"""Committed fixtures (tests/golden/golden.npz, written by gen_golden.py from
the oracle) reproduced by the oracle (CPU) and by the HIP path (GPU)."""
import os

import numpy as np
import pytest

from conftest import ROOT, keys_for, oracle_for


def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"))


def test_chacha_stream_fixture():
    from oracle.oracle import lib, seed_bytes
    L = lib()
    r = L.pzo_rng_new(seed_bytes(0))
    got = np.array([L.pzo_rng_next_u64(r) for _ in range(16)], dtype=np.uint64)
    L.pzo_rng_free(r)
    assert np.array_equal(got, golden()["chacha8_seed0_u64"])


@pytest.mark.parametrize("name", ["SP_TEST", "NI_2P"])
def test_oracle_reproduces_fixtures(name):
    import hashlib
    G = golden()
    o, s, z, key = keys_for(name)
    assert o.psi() == int(G[name + "_psi"][0])
    x = G[name + "_ntt_in"]
    for which in ("forward", "forward_lazy", "backward", "backward_lazy"):
        assert np.array_equal(o.ntt_rows(which, x), G[name + "_ntt_" + which])
    h = hashlib.sha256()
    for a in (key.rgsw, key.auto_key, key.lwe_ksk, s, z):
        h.update(np.ascontiguousarray(a).tobytes())
    assert np.array_equal(np.frombuffer(h.digest(), dtype=np.uint8), G[name + "_key_digest"])
    rg, rgs = key.rgsw_ct(3)
    assert np.array_equal(o.rlwe_by_rgsw_shoup(G[name + "_ext_in"], rg, rgs, False),
                          G[name + "_ext_out"])
    kt, kts = key.auto_ct(2)
    assert np.array_equal(o.rlwe_auto_shoup(G[name + "_ext_in"], kt, kts, 2, False),
                          G[name + "_auto_out"])
    ops = G[name + "_gate_ops"]
    for i in range(4):
        out = o.gate(key, int(ops[i]), G[name + "_gate_c0"][i], G[name + "_gate_c1"][i])
        assert np.array_equal(out, G[name + "_gate_out"][i])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["SP_TEST", "NI_2P"])
def test_gpu_reproduces_fixtures(name):
    import pzlib
    G = golden()
    o, s, z, key = keys_for(name)
    ctx = pzlib.Context(name)
    print("golden gpu %s" % name, flush=True)
    for which in ("forward", "forward_lazy", "backward", "backward_lazy"):
        d = ctx.to_device(G[name + "_ntt_in"])
        ctx.ntt(which, d, 2)
        assert np.array_equal(d.download(), G[name + "_ntt_" + which])
    k = ctx.key_upload(key.rgsw, key.auto_key, key.lwe_ksk)
    d = ctx.to_device(G[name + "_ext_in"][None])
    ctx.rlwe_by_rgsw(k, d, ctx.to_device(np.array([3], dtype=np.uint32)),
                     ctx.to_device(np.array([0], dtype=np.uint8)), 1)
    assert np.array_equal(d.download()[0], G[name + "_ext_out"])
    d = ctx.to_device(G[name + "_ext_in"][None])
    ctx.rlwe_auto(k, d, ctx.to_device(np.array([2], dtype=np.uint32)),
                  ctx.to_device(np.array([0], dtype=np.uint8)), 1)
    assert np.array_equal(d.download()[0], G[name + "_auto_out"])
    dout = ctx.array((4, o.N + 1))
    ctx.gates(k, ctx.to_device(G[name + "_gate_ops"]), ctx.to_device(G[name + "_gate_c0"]),
              ctx.to_device(G[name + "_gate_c1"]), dout, 4)
    assert np.array_equal(dout.download(), G[name + "_gate_out"])
    k.free()
    ctx.close()
