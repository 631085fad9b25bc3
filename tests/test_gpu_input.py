"""GPU: the input path of the non-interactive protocol against the oracle.

Party secret, Ksk(u -> s) and the batched RLWE bit encryption are byte-exact
with the oracle's seeded procedures; rlwe_key_switch and sample_extract are
bit-exact on random inputs (also in place); end to end, bits encrypted under
a party's u, switched and extracted on the GPU decrypt under the ideal s and
feed the gate bootstraps."""
import sys

import numpy as np
import pytest

from conftest import keys_for, rng, uniform_mod

pytestmark = pytest.mark.gpu
_ctx = {}


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def ctx_for(name):
    import pzlib
    if name not in _ctx:
        _ctx[name] = pzlib.Context(name)
    return _ctx[name]


@pytest.mark.parametrize("name", ["NI_2P", "NI_8P"])
def test_keys_and_encryption_match_oracle(name):
    o, s, z, key = keys_for(name)
    c = ctx_for(name)
    log("input keys %s" % name)
    u = c.keygen_user_secret(17)
    assert np.array_equal(u, o.keygen_user_secret(17))
    k = c.keygen_rlwe_ksk(18, u, s)
    assert np.array_equal(k.download(), o.keygen_rlwe_ksk(18, u, s))
    k.free()
    bits = rng(310).integers(0, 2, size=c.N + 5).astype(np.uint8)
    assert np.array_equal(c.encrypt_rlwe_bools(u, bits, 19), o.encrypt_rlwe_bools(u, bits, 19))


def test_rlwe_key_switch_parity():
    o, s, z, key = keys_for("NI_2P")
    c = ctx_for("NI_2P")
    B, d = 24, o.p.uitos_d
    log("rlwe key switch B=%d" % B)
    r = rng(311)
    x = uniform_mod(r, o.Q, (B, 2, o.N))
    x[0, 0, :4] = [0, 1, o.Q - 1, o.Q // 2]
    rows = uniform_mod(r, o.Q, (2 * d, o.N))
    rows_s = o.shoup_rows(rows)
    k = c.rlwe_ksk_upload(rows)
    want = np.stack([o.rlwe_key_switch(x[b], rows, rows_s) for b in range(B)])
    din = c.to_device(x)
    dout = c.array((B, 2, o.N))
    c.rlwe_key_switch(k, din, dout, B)
    assert np.array_equal(dout.download(), want)
    c.rlwe_key_switch(k, din, din, B)  # in place
    assert np.array_equal(din.download(), want)
    k.free()


def test_sample_extract_parity():
    o, s, z, key = keys_for("NI_2P")
    c = ctx_for("NI_2P")
    r = rng(312)
    R = 3
    x = uniform_mod(r, o.Q, (R, 2, o.N))
    x[1, 0, :8] = 0  # neg(0) = q where the extraction negates
    idx = [(0, 0), (0, 1), (1, o.N - 1), (2, 777), (1, 3), (2, 0)]
    src = np.array(idx, dtype=np.uint32)
    log("sample extract count=%d" % len(idx))
    want = np.stack([o.sample_extract(x[ri], k) for ri, k in idx])
    dl = c.array((len(idx), o.N + 1))
    c.sample_extract(c.to_device(x), c.to_device(src), len(idx), dl)
    assert np.array_equal(dl.download(), want)


@pytest.mark.parametrize("name", ["NI_2P", "NI_8P"])
def test_input_path_end_to_end(name):
    o, s, z, key = keys_for(name)
    c = ctx_for(name)
    log("input path end to end %s" % name)
    u = c.keygen_user_secret(21)
    k = c.keygen_rlwe_ksk(22, u, s)
    bits = rng(313).integers(0, 2, size=2 * c.N + 100).astype(np.uint8)
    cts = c.encrypt_rlwe_bools(u, bits, 23)
    R, count = cts.shape[0], len(bits)
    d = c.to_device(cts)
    c.rlwe_key_switch(k, d, d, R)
    src = np.array([(t // c.N, t % c.N) for t in range(count)], dtype=np.uint32)
    dl = c.array((count, c.N + 1))
    c.sample_extract(d, c.to_device(src), count, dl)
    lwes = dl.download()
    assert np.array_equal(c.decrypt_bools(s, lwes), bits)
    # the extracted inputs feed gate bootstraps: NAND of consecutive pairs
    B = 64
    dk = c.key_upload(key.rgsw, key.auto_key, key.lwe_ksk)
    d0, d1 = c.to_device(lwes[0:2 * B:2]), c.to_device(lwes[1:2 * B:2])
    dout = c.array((B, c.N + 1))
    c.gates(dk, c.to_device(np.zeros(B, dtype=np.uint32)), d0, d1, dout, B)
    want = (1 - (bits[0:2 * B:2] & bits[1:2 * B:2])).astype(np.uint8)
    assert np.array_equal(c.decrypt_bools(s, dout.download()), want)
    dk.free()
    k.free()
