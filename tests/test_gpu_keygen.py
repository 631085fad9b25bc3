"""GPU: the product's seeded key generation and client encryption produce
exactly the oracle's bytes (same seeds), and gates under a product key
decrypt correctly."""
import numpy as np
import pytest

from conftest import keys_for, oracle_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["SP_TEST", "NI_2P"])
def test_keygen_matches_oracle(name):
    import pzlib
    print("keygen %s" % name, flush=True)
    o, s, z, key = keys_for(name, 0, 1, 2)
    ctx = pzlib.Context(name)
    ps, pzv = ctx.keygen_secrets(0, 2)
    assert np.array_equal(ps, s) and np.array_equal(pzv, z)
    k = ctx.keygen_server(1, ps, pzv)
    rgsw, auto, ksk = ctx.key_download(k)
    assert np.array_equal(rgsw.reshape(-1), key.rgsw)
    assert np.array_equal(auto.reshape(-1), key.auto_key)
    assert np.array_equal(ksk.reshape(-1), key.lwe_ksk)
    k.free()
    ctx.close()


@pytest.mark.parametrize("name", ["SP_TEST", "NI_2P"])
def test_encrypt_matches_oracle_and_gates_decrypt(name):
    import pzlib
    print("encrypt %s" % name, flush=True)
    o, s, z, key = keys_for(name, 0, 1, 2)
    ctx = pzlib.Context(name)
    bits0 = np.array([0, 0, 1, 1, 0, 1], dtype=np.uint8)
    bits1 = np.array([0, 1, 0, 1, 1, 1], dtype=np.uint8)
    c0 = ctx.encrypt_bools(s, bits0, 7000)
    c1 = ctx.encrypt_bools(s, bits1, 8000)
    for i in range(len(bits0)):
        assert np.array_equal(c0[i], o.encrypt_bool(s, int(bits0[i]), 7000 + i))
    assert np.array_equal(ctx.decrypt_bools(s, c0), bits0)
    k = ctx.keygen_server(1, s, z)
    ops = np.array([0, 1, 2, 3, 4, 5], dtype=np.uint32)
    dout = ctx.array((6, o.N + 1))
    ctx.gates(k, ctx.to_device(ops), ctx.to_device(c0), ctx.to_device(c1), dout, 6)
    got = ctx.decrypt_bools(s, dout.download())
    want = [1 - (bits0[0] & bits1[0]), bits0[1] & bits1[1], bits0[2] | bits1[2],
            1 - (bits0[3] | bits1[3]), bits0[4] ^ bits1[4], 1 - (bits0[5] ^ bits1[5])]
    assert [int(v) for v in got] == [int(v) for v in want]
    k.free()
    ctx.close()
