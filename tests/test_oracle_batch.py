"""CPU: the oracle's threaded gate batch gives the same words as one gate at
a time, and the cone evaluation used by the circuit tests equals full serial
evaluation of a small FheUint8 DAG."""
import numpy as np

from conftest import keys_for, oracle_cone


def test_gates_batch_equals_gate():
    o, s, z, key = keys_for("SP_TEST")
    r = np.random.default_rng(3)
    count = 12
    gates = r.integers(0, 6, size=count).astype(np.uint32)
    b0 = r.integers(0, 2, size=count)
    b1 = r.integers(0, 2, size=count)
    c0 = np.stack([o.encrypt_bool(s, int(b), 50 + i) for i, b in enumerate(b0)])
    c1 = np.stack([o.encrypt_bool(s, int(b), 90 + i) for i, b in enumerate(b1)])
    got = o.gates_batch(key, gates, c0, c1, threads=4)
    for i in range(count):
        assert np.array_equal(got[i], o.gate(key, int(gates[i]), c0[i], c1[i])), i
    assert np.array_equal(c0[0], o.encrypt_bool(s, int(b0[0]), 50))  # inputs untouched


def test_cone_equals_serial():
    import pzlib
    o, s, z, key = keys_for("SP_TEST")
    circ = pzlib.Circuit(None)
    wa, wb = circ.inputs(8), circ.inputs(8)
    lt = circ.u8("lt", wa, wb)
    nt = circ.gate("not", int(lt[0]))
    a, b = 23, 200
    bits = [(a >> i) & 1 for i in range(8)] + [(b >> i) & 1 for i in range(8)]
    cts = np.stack([o.encrypt_bool(s, v, 700 + i) for i, v in enumerate(bits)])
    nodes = circ.nodes()
    cone = oracle_cone(o, key, nodes, cts, [int(lt[0]), nt], threads=4)
    vals, k = {}, 0
    for i, (kind, op, x, y) in enumerate(nodes):
        if kind == 0:
            vals[i] = cts[k]
            k += 1
        elif kind == 1:
            vals[i] = o.gate(key, int(op), vals[int(x)], vals[int(y)])
        else:
            vals[i] = o.not_(vals[int(x)])
    assert np.array_equal(cone[int(lt[0])], vals[int(lt[0])])
    assert np.array_equal(cone[nt], vals[nt])
    assert o.decrypt_bool(s, cone[int(lt[0])]) == int(a < b)
    circ.close()
