"""Random FheUint8 programs for the serving loop and the benchmarks.

A program has `n_in` encrypted u8 inputs and `n_ops` operations; each
operation draws its two operands from the inputs or earlier results and is
one of add / sub / mul / div / rem / lt / eq.  `plain` evaluates the same
program on integers (FheUint8 semantics: mod-256 arithmetic, x / 0 = 255,
x % 0 = x).
"""
import numpy as np

ARITH = ("add", "sub", "mul", "div", "rem")
COMPARE = ("lt", "eq")


def plain_op(op, a, b):
    if op == "add":
        return (a + b) & 255
    if op == "sub":
        return (a - b) & 255
    if op == "mul":
        return (a * b) & 255
    if op == "div":
        return a // b if b else 255
    if op == "rem":
        return a % b if b else a
    if op == "lt":
        return int(a < b)
    if op == "eq":
        return int(a == b)
    raise ValueError(op)


class Program:
    def __init__(self, rng, n_ops, n_in=2, ops=ARITH + COMPARE, p_input=0.5):
        self.n_in = n_in
        self.values = [int(v) for v in rng.integers(0, 256, size=n_in)]
        self.steps = []  # (op, src_a, src_b) over value indices
        n_vals = n_in
        for _ in range(n_ops):
            op = ops[int(rng.integers(0, len(ops)))]
            pick = lambda: int(rng.integers(0, n_in)) if (rng.random() < p_input or n_vals == n_in) \
                else int(rng.integers(n_in, n_vals))  # noqa: E731
            a, b = pick(), pick()
            self.steps.append((op, a, b))
            if op in ARITH:
                n_vals += 1

    def plain(self):
        vals = list(self.values)
        bools = []
        for op, a, b in self.steps:
            r = plain_op(op, vals[a], vals[b])
            (vals.append(r) if op in ARITH else bools.append(r))
        return vals[self.n_in:], bools

    def build(self, circuit):
        """Adds the program to a pzlib.Circuit; returns (output wires,
        number of u8 results, number of bool results)."""
        wires = [circuit.inputs(8) for _ in range(self.n_in)]
        u8_out, bool_out = [], []
        for op, a, b in self.steps:
            out = circuit.u8(op, wires[a], wires[b])
            if op in ARITH:
                wires.append(out)
                u8_out.append(out)
            else:
                bool_out.append(out[0])
        flat = [int(w) for o in u8_out for w in o] + [int(w) for w in bool_out]
        return np.array(flat, dtype=np.uint32), len(u8_out), len(bool_out)

    def input_bits(self):
        return np.array([(v >> i) & 1 for v in self.values for i in range(8)], dtype=np.uint8)


def decode(bits, n_u8, n_bool):
    u8 = [sum(int(bits[8 * k + i]) << i for i in range(8)) for k in range(n_u8)]
    return u8, [int(b) for b in bits[8 * n_u8:8 * n_u8 + n_bool]]
