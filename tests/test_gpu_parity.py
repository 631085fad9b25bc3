"""GPU parity: every operator of the C ABI against the oracle, bit for bit.

Each test runs the HIP path (libpz.so through ctypes) and the oracle on the
same seeded inputs and keys and compares the full output arrays.  Sizes are
chosen so the oracle side finishes in seconds.
"""
import sys

import numpy as np
import pytest

from conftest import keys_for, oracle_for, rng, uniform_mod

pytestmark = pytest.mark.gpu

# NI_4P_HB_FR: br_q = 2048 (embedding factor 2); NI_8P: q_ks = 2^17 (32-bit KSK)
PARAMS = ["NI_2P", "SP_TEST", "NI_4P_HB_FR", "NI_8P"]
# blind-rotation kernels: 1 = one work group per bootstrap, 2 = two half
# groups, 3 = two half groups with 4 coefficients per thread (2 and 3: d = 1
# sets only); all must give the same words
MODES = [(p, m) for p in PARAMS for m in ((1,) if p == "SP_TEST" else (1, 2, 3))]
_ctx = {}
_dkeys = {}


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def ctx_for(name):
    import pzlib
    if name not in _ctx:
        _ctx[name] = pzlib.Context(name)
    return _ctx[name]


def dkey_for(name):
    if name not in _dkeys:
        o, s, z, key = keys_for(name)
        c = ctx_for(name)
        _dkeys[name] = c.key_upload(key.rgsw, key.auto_key, key.lwe_ksk)
    return _dkeys[name]


@pytest.mark.parametrize("pname", PARAMS)
def test_ntt_tables_match(pname):
    log("ntt tables %s" % pname)
    o = oracle_for(pname)
    c = ctx_for(pname)
    assert np.array_equal(c.ntt_tables(), o.ntt_tables())
    assert c.info()["psi"] == o.psi()


@pytest.mark.parametrize("pname", PARAMS)
@pytest.mark.parametrize("which", ["forward", "forward_lazy", "backward", "backward_lazy"])
def test_ntt_rows(pname, which):
    o = oracle_for(pname)
    c = ctx_for(pname)
    rows = 512 if o.N == 2048 else 2048
    log("ntt %s %s rows=%d" % (pname, which, rows))
    r = rng(11)
    # inputs in the reference's accepted ranges: [0,2q) for lazy entry points
    hi = 2 * o.Q if which.startswith("backward") or which == "forward_lazy" else o.Q
    x = uniform_mod(r, hi, (rows, o.N))
    x[0, :] = 0
    x[1, :] = hi - 1
    d = c.to_device(x)
    c.ntt(which, d, rows)
    got = d.download()
    want = o.ntt_rows(which, x)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("pname", PARAMS)
def test_shoup_matrix_fma(pname):
    o = oracle_for(pname)
    c = ctx_for(pname)
    log("fma %s" % pname)
    import oracle.oracle as om
    L = om.lib()
    r = rng(12)
    d, n = 3, 4 * o.N
    a = uniform_mod(r, o.Q, (d, n))
    a_s = np.array([[L.pzo_shoup_repr(int(v), o.Q) for v in row] for row in a], dtype=np.uint64)
    b = r.integers(0, 2 ** 63, size=(d, n), dtype=np.uint64) * 2 + 1
    out = uniform_mod(r, 2 * o.Q, n)
    want = out.copy()
    L.pzo_shoup_matrix_fma(o.Q, want, a.reshape(-1), a_s.reshape(-1), b.reshape(-1), d, n)
    dout = c.to_device(out)
    c.shoup_matrix_fma(dout, c.to_device(a), c.to_device(a_s), c.to_device(b), d, n)
    assert np.array_equal(dout.download(), want)


@pytest.mark.parametrize("pname", PARAMS)
@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_decompose(pname, which):
    o = oracle_for(pname)
    c = ctx_for(pname)
    log("decompose %s %d" % (pname, which))
    import oracle.oracle as om
    L = om.lib()
    p = o.p
    q, logb, d = [(p.rlwe_q, p.rlrg_logb, p.rlrg_da), (p.rlwe_q, p.rlrg_logb, p.rlrg_db),
                  (p.rlwe_q, p.auto_logb, p.auto_d), (p.lwe_q, p.lwe_logb, p.lwe_d)][which]
    r = rng(13 + which)
    vals = uniform_mod(r, q, 20000)
    edge = np.array([0, 1, q - 1, q // 2, q // 2 - 1, q // 2 + 1, q], dtype=np.uint64)
    vals = np.concatenate([vals, edge])
    want = np.zeros(len(vals) * d, dtype=np.uint64)
    L.pzo_decompose_iter(q, logb, d, vals, want, len(vals))
    want = want.reshape(len(vals), d).T.copy()
    dd = c.array((d, len(vals)))
    c.decompose(which, c.to_device(vals), dd, len(vals))
    assert np.array_equal(dd.download(), want)


def _rlwe_batch(o, B, seed):
    r = rng(seed)
    return uniform_mod(r, o.Q, (B, 2, o.N))


@pytest.mark.parametrize("pname", PARAMS)
def test_rlwe_by_rgsw(pname):
    o, s, z, key = keys_for(pname)
    c = ctx_for(pname)
    k = dkey_for(pname)
    B = 64
    log("rlwe_by_rgsw %s B=%d" % (pname, B))
    x = _rlwe_batch(o, B, 21)
    r = rng(22)
    idx = r.integers(0, o.n, size=B, dtype=np.uint32)
    triv = (np.arange(B) % 4 == 0).astype(np.uint8)
    x[triv == 1, 0, :] = 0
    want = np.empty_like(x)
    for b in range(B):
        ct, cts = key.rgsw_ct(int(idx[b]))
        want[b] = o.rlwe_by_rgsw_shoup(x[b], ct, cts, bool(triv[b]))
    d = c.to_device(x)
    c.rlwe_by_rgsw(k, d, c.to_device(idx), c.to_device(triv), B)
    assert np.array_equal(d.download(), want)


@pytest.mark.parametrize("pname", PARAMS)
def test_rlwe_auto(pname):
    o, s, z, key = keys_for(pname)
    c = ctx_for(pname)
    k = dkey_for(pname)
    B = 2 * (o.p.w + 1)
    log("rlwe_auto %s B=%d" % (pname, B))
    x = _rlwe_batch(o, B, 31)
    kidx = np.array([i % (o.p.w + 1) for i in range(B)], dtype=np.uint32)
    triv = (np.arange(B) >= o.p.w + 1).astype(np.uint8)
    x[triv == 1, 0, :] = 0
    x[-1, 1, :5] = 0  # zero coefficients exercise neg(0) = q in the trivial path
    want = np.empty_like(x)
    for b in range(B):
        kt, kts = key.auto_ct(int(kidx[b]))
        want[b] = o.rlwe_auto_shoup(x[b], kt, kts, int(kidx[b]), bool(triv[b]))
    d = c.to_device(x)
    c.rlwe_auto(k, d, c.to_device(kidx), c.to_device(triv), B)
    assert np.array_equal(d.download(), want)


@pytest.mark.parametrize("pname", PARAMS)
def test_lwe_key_switch(pname):
    o, s, z, key = keys_for(pname)
    c = ctx_for(pname)
    k = dkey_for(pname)
    B = 40
    log("lwe_key_switch %s B=%d" % (pname, B))
    r = rng(41)
    x = r.integers(0, o.p.lwe_q, size=(B, o.N + 1), dtype=np.uint64)
    x[0, :] = o.p.lwe_q      # the moddown edge value (wraps to digit 0)
    x[1, :] = o.p.lwe_q - 1
    x[2, :] = 0
    want = np.stack([o.lwe_key_switch(x[b], key.lwe_ksk) for b in range(B)])
    dout = c.array((B, o.n + 1))
    c.lwe_key_switch(k, c.to_device(x), dout, B)
    assert np.array_equal(dout.download(), want)


def _gate_inputs(o, s, B, seed):
    r = rng(seed)
    bits0 = r.integers(0, 2, size=B)
    bits1 = r.integers(0, 2, size=B)
    c0 = np.stack([o.encrypt_bool(s, int(bits0[i]), seed * 100000 + 2 * i) for i in range(B)])
    c1 = np.stack([o.encrypt_bool(s, int(bits1[i]), seed * 100000 + 2 * i + 1) for i in range(B)])
    return bits0, bits1, c0, c1


PLAIN = {0: lambda a, b: 1 - (a & b), 1: lambda a, b: a & b, 2: lambda a, b: a | b,
         3: lambda a, b: 1 - (a | b), 4: lambda a, b: a ^ b, 5: lambda a, b: 1 - (a ^ b)}


@pytest.mark.parametrize("pname,mode", MODES)
def test_pbs(pname, mode):
    o, s, z, key = keys_for(pname)
    c = ctx_for(pname)
    k = dkey_for(pname)
    B = 6 if o.N == 2048 else 48
    log("pbs %s mode=%d B=%d" % (pname, mode, B))
    c.set_blind_rotate_mode(mode)
    r = rng(51)
    x = uniform_mod(r, o.Q, (B, o.N + 1))
    x[0, :8] = o.Q - 1  # coefficients whose f64 moddown rounds to q_ks
    gate = (np.arange(B) % 6).astype(np.uint32)
    want = np.stack([o.pbs(key, int(gate[b]), x[b])[0] for b in range(B)])
    d = c.to_device(x)
    c.pbs(k, c.to_device(gate), d, B)
    c.set_blind_rotate_mode(0)
    assert np.array_equal(d.download(), want)


@pytest.mark.parametrize("pname,mode", MODES)
def test_pbs_degenerate_buckets(pname, mode):
    """Key-switch outputs whose dlog buckets are degenerate: a zero and a
    trivial ciphertext (every LWE index in one bucket, the op list's longest
    bucket), ciphertexts with one or a few non-zero mask coefficients (long
    runs of empty buckets, forced automorphisms every w buckets), next to a
    uniform one."""
    o, s, z, key = keys_for(pname)
    c = ctx_for(pname)
    k = dkey_for(pname)
    log("pbs degenerate %s mode=%d" % (pname, mode))
    c.set_blind_rotate_mode(mode)
    r = rng(53)
    x = np.zeros((6, o.N + 1), dtype=np.uint64)
    x[1, 0] = o.Q // 8                       # trivial: b only
    x[2, 0], x[2, 5] = o.Q // 3, o.Q // 5    # one mask coefficient
    x[3, 1:9] = o.Q - 1                      # a few, all rounding to q_ks
    x[4, 1::97] = uniform_mod(r, o.Q, (len(range(1, o.N + 1, 97)),))
    x[5] = uniform_mod(r, o.Q, (o.N + 1,))
    B = x.shape[0]
    gate = (np.arange(B) % 6).astype(np.uint32)
    want = np.stack([o.pbs(key, int(gate[b]), x[b])[0] for b in range(B)])
    d = c.to_device(x)
    c.pbs(k, c.to_device(gate), d, B)
    c.set_blind_rotate_mode(0)
    assert np.array_equal(d.download(), want)


@pytest.mark.parametrize("pname,mode", MODES)
def test_gates_truth_and_parity(pname, mode):
    o, s, z, key = keys_for(pname)
    c = ctx_for(pname)
    k = dkey_for(pname)
    B = 12 if o.N == 2048 else 96
    log("gates %s mode=%d B=%d" % (pname, mode, B))
    c.set_blind_rotate_mode(mode)
    bits0, bits1, c0, c1 = _gate_inputs(o, s, B, 61)
    ops = (np.arange(B) % 6).astype(np.uint32)
    want = np.stack([o.gate(key, int(ops[b]), c0[b], c1[b]) for b in range(B)])
    d0, d1, dops = c.to_device(c0), c.to_device(c1), c.to_device(ops)
    dout = c.array((B, o.N + 1))
    for rep in range(3):  # first call captures the graph, later calls replay it
        c.gates(k, dops, d0, d1, dout, B)
        got = dout.download()
        assert np.array_equal(got, want), "replay %d" % rep
    c.set_blind_rotate_mode(0)
    for b in range(B):
        assert o.decrypt_bool(s, got[b]) == PLAIN[int(ops[b])](int(bits0[b]), int(bits1[b]))


def test_split_mode_needs_single_digit_decompositions():
    import pzlib
    c = ctx_for("SP_TEST")
    for mode in (2, 3):
        with pytest.raises(pzlib.PzError):
            c.set_blind_rotate_mode(mode)
    c.set_blind_rotate_mode(0)


# The reference's other parameter sets (src/bool/parameters.rs:526-596,
# :652-677): q_ks = 2^15 with 12 key-switch digits (I_2P_LB_SR), n = 580 /
# 620 / 660, br_q = 2048 and 4096, the interactive sets' ternary LWE secret.
OTHER_SETS = ["I_2P_LB_SR", "I_4P", "I_8P", "NI_4P_LB_SR"]


@pytest.mark.parametrize("pname", OTHER_SETS)
def test_pbs_and_gates_other_sets(pname):
    o, s, z, key = keys_for(pname)
    c = ctx_for(pname)
    k = dkey_for(pname)
    B = 6
    log("pbs + gates %s B=%d" % (pname, B))
    r = rng(71)
    x = uniform_mod(r, o.Q, (B, o.N + 1))
    x[0, :8] = o.Q - 1
    gate = (np.arange(B) % 6).astype(np.uint32)
    want = np.stack([o.pbs(key, int(gate[b]), x[b])[0] for b in range(B)])
    d = c.to_device(x)
    c.pbs(k, c.to_device(gate), d, B)
    assert np.array_equal(d.download(), want)
    bits0, bits1, c0, c1 = _gate_inputs(o, s, B, 72)
    ops = (np.arange(B) % 6).astype(np.uint32)
    want = np.stack([o.gate(key, int(ops[b]), c0[b], c1[b]) for b in range(B)])
    dout = c.array((B, o.N + 1))
    c.gates(k, c.to_device(ops), c.to_device(c0), c.to_device(c1), dout, B)
    got = dout.download()
    assert np.array_equal(got, want)
    for b in range(B):
        assert o.decrypt_bool(s, got[b]) == PLAIN[int(ops[b])](int(bits0[b]), int(bits1[b]))
