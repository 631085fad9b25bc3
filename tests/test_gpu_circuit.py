"""GPU: FheUint8 circuits evaluated level by level (every ready gate of a
level batched into one launch) decrypt to plaintext u8 semantics, and each
output ciphertext equals the oracle's serial evaluation of the same gate
DAG, gate by gate in the reference's order."""
import numpy as np
import pytest

from conftest import keys_for, oracle_cone

pytestmark = pytest.mark.gpu

OPS = {
    "add": lambda a, b: [(a + b) & 255], "sub": lambda a, b: [(a - b) & 255],
    "mul": lambda a, b: [(a * b) & 255],
    "div_rem": lambda a, b: [a // b if b else 255, a % b if b else a],
    "lt": lambda a, b: [int(a < b)], "gt": lambda a, b: [int(a > b)],
    "le": lambda a, b: [int(a <= b)], "ge": lambda a, b: [int(a >= b)],
    "eq": lambda a, b: [int(a == b)], "neq": lambda a, b: [int(a != b)],
    "max": lambda a, b: [max(a, b)], "min": lambda a, b: [min(a, b)],
    "overflowing_add": lambda a, b: [(a + b) & 255, int(a + b > 255)],
    "overflowing_sub": lambda a, b: [(a - b) & 255, int(a < b)],
}


def bits(v):
    return [(v >> i) & 1 for i in range(8)]


def unpack(op, out_bits):
    if op in ("lt", "gt", "le", "ge", "eq", "neq"):
        return [int(out_bits[0])]
    vals = [sum(int(out_bits[8 * k + i]) << i for i in range(8)) for k in range(len(out_bits) // 8)]
    if op.startswith("overflowing"):
        vals.append(int(out_bits[8]))
    return vals


def serial_oracle(o, key, nodes, inputs_ct):
    """Evaluate the DAG gate by gate in node order with the oracle."""
    vals = {}
    k = 0
    for i, (kind, op, a, b) in enumerate(nodes):
        if kind == 0:
            vals[i] = inputs_ct[k]
            k += 1
        elif kind == 1:
            vals[i] = o.gate(key, int(op), vals[int(a)], vals[int(b)])
        else:
            vals[i] = o.not_(vals[int(a)])
    return vals


def run_case(name, ops, pairs, seed, check_serial):
    import pzlib
    o, s, z, key = keys_for(name)
    ctx = pzlib.Context(name)
    k = ctx.key_upload(key.rgsw, key.auto_key, key.lwe_ksk)
    circ = pzlib.Circuit(ctx)
    in_bits, outs, specs = [], [], []
    for op, (a, b) in zip(ops, pairs):
        wa, wb = circ.inputs(8), circ.inputs(8)
        in_bits += bits(a) + bits(b)
        sel = None
        outs.append(circ.u8(op, wa, wb, sel))
        specs.append((op, a, b))
    cts = ctx.encrypt_bools(s, np.array(in_bits, dtype=np.uint8), seed)
    print("circuit %s ops=%d pbs=%d" % (name, len(ops), circ.stats()["pbs"]), flush=True)
    circ.run(k, ctx.to_device(cts))
    all_out = np.concatenate(outs)
    got_ct = circ.read(all_out)
    got_bits = ctx.decrypt_bools(s, got_ct)
    pos = 0
    for (op, a, b), w in zip(specs, outs):
        assert unpack(op, got_bits[pos:pos + len(w)]) == OPS[op](a, b), (op, a, b)
        pos += len(w)
    if check_serial:
        vals = serial_oracle(o, key, circ.nodes(), cts)
        for i, wire in enumerate(all_out):
            assert np.array_equal(got_ct[i], vals[int(wire)]), ("wire", int(wire))
    circ.close()
    k.free()
    ctx.close()


def test_u8_ops_sp_test_bit_exact_vs_serial():
    r = np.random.default_rng(90)
    ops = sorted(OPS)
    pairs = [(int(r.integers(0, 256)), int(r.integers(0, 256))) for _ in ops]
    pairs[ops.index("div_rem")] = (200, 0)  # division by zero: quotient 255, remainder a
    run_case("SP_TEST", ops, pairs, 12345, check_serial=True)


def test_u8_mul_lt_div_ni2p():
    r = np.random.default_rng(91)
    ops = ["mul", "lt", "div_rem", "mul"]
    pairs = [(int(r.integers(0, 256)), int(r.integers(1, 256))) for _ in ops]
    run_case("NI_2P", ops, pairs, 777, check_serial=False)


def test_circuit_grows_between_runs_sp_test():
    """A circuit run, extended (a new operation on its outputs and a chain of
    NOTs, which run in dependent waves of one level) and run again: the second
    run uses the new schedule, and every output equals the oracle's serial
    evaluation of the grown DAG."""
    import pzlib
    o, s, z, key = keys_for("SP_TEST")
    ctx = pzlib.Context("SP_TEST")
    k = ctx.key_upload(key.rgsw, key.auto_key, key.lwe_ksk)
    circ = pzlib.Circuit(ctx)
    a, b = 201, 77
    wa, wb = circ.inputs(8), circ.inputs(8)
    s1 = circ.u8("add", wa, wb)
    cts = ctx.encrypt_bools(s, np.array(bits(a) + bits(b), dtype=np.uint8), 4242)
    d_in = ctx.to_device(cts)
    circ.run(k, d_in)
    got = ctx.decrypt_bools(s, circ.read(s1))
    assert unpack("add", got) == OPS["add"](a, b)
    # grow: (a + b) * b, and NOT(NOT(NOT(bit 0 of the sum)))
    p = circ.u8("mul", s1, wb)
    n1 = circ.gate("not", int(s1[0]))
    n2 = circ.gate("not", n1)
    n3 = circ.gate("not", n2)
    print("grown circuit pbs=%d" % circ.stats()["pbs"], flush=True)
    circ.run(k, d_in)
    wires = np.concatenate([s1, p, np.array([n1, n2, n3], dtype=np.uint32)])
    got_ct = circ.read(wires)
    got = ctx.decrypt_bools(s, got_ct)
    assert unpack("add", got[:8]) == OPS["add"](a, b)
    assert unpack("mul", got[8:16]) == OPS["mul"]((a + b) & 255, b)
    bit0 = ((a + b) & 255) & 1
    assert [int(x) for x in got[16:19]] == [1 - bit0, bit0, 1 - bit0]
    vals = serial_oracle(o, key, circ.nodes(), cts)
    for i, wire in enumerate(wires):
        assert np.array_equal(got_ct[i], vals[int(wire)]), ("wire", int(wire))
    circ.close()
    k.free()
    ctx.close()


def test_u8_mul_divrem_lt_ni8p():
    """FheUint8 at NI_8P (8-party keys, q_ks = 2^17: 32-bit key-switch
    batches at circuit scale): 16 mul + 16 div_rem (one by zero) + 16 lt in
    one run.  Every output decrypts to u8 arithmetic, the div-by-zero flag
    decrypts to 1, and every wire of one mul and one lt equals the oracle's
    evaluation of the same gates."""
    import pzlib
    name = "NI_8P"
    o, s, z, key = keys_for(name)
    ctx = pzlib.Context(name)
    k = ctx.key_upload(key.rgsw, key.auto_key, key.lwe_ksk)
    circ = pzlib.Circuit(ctx)
    r = np.random.default_rng(92)
    ops = ["mul"] * 16 + ["div_rem"] * 16 + ["lt"] * 16
    pairs = [(int(r.integers(0, 256)), int(r.integers(0, 256))) for _ in ops]
    pairs[16] = (173, 0)  # division by zero: quotient 255, remainder a
    in_bits, outs = [], []
    for op, (a, b) in zip(ops, pairs):
        wa, wb = circ.inputs(8), circ.inputs(8)
        in_bits += bits(a) + bits(b)
        outs.append(circ.u8(op, wa, wb))
    flag = circ.div_zero_flag()
    st = circ.stats()
    print("circuit NI_8P ops=%d pbs=%d depth=%d widest=%d" %
          (len(ops), st["pbs"], st["depth"], int(circ.level_sizes().max())), flush=True)
    cts = ctx.encrypt_bools(s, np.array(in_bits, dtype=np.uint8), 4321)
    circ.run(k, ctx.to_device(cts))
    all_out = np.concatenate(outs + [np.array([flag], dtype=np.uint32)])
    got_ct = circ.read(all_out)
    got_bits = ctx.decrypt_bools(s, got_ct)
    pos = 0
    for op, (a, b), w in zip(ops, pairs, outs):
        assert unpack(op, got_bits[pos:pos + len(w)]) == OPS[op](a, b), (op, a, b)
        pos += len(w)
    assert int(got_bits[-1]) == 1  # the div-by-zero flag
    # one mul (index 3) and one lt (index 40): every output wire vs the oracle
    check = [int(w) for w in outs[3]] + [int(outs[40][0])]
    print("oracle cone NI_8P wires=%d" % len(check), flush=True)
    want = oracle_cone(o, key, circ.nodes(), cts, check)
    pos_of = {int(w): i for i, w in enumerate(all_out)}
    for w in check:
        assert np.array_equal(got_ct[pos_of[w]], want[w]), ("wire", w)
    circ.close()
    k.free()
    ctx.close()
