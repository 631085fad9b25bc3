"""FheUint8 circuit structure (no GPU): gate counts and depths of every
operation, built from the C++ restatement of src/shortint/ops.rs.

Expected figures come from tracing the reference's gate calls: mul 155/24,
div_rem 307/109 (including the 7-gate is_zero of the div-by-zero flag), lt
35/9, add 37/15, sub 38/16, eq 15/8, mux 24/2 (PBS / depth)."""
import pytest

EXPECTED = {"add": (37, 15), "sub": (38, 16), "mul": (155, 24), "div_rem": (307, 109),
            "lt": (35, 9), "eq": (15, 8), "mux": (24, 2), "gt": (35, 9), "le": (35, 9),
            "ge": (35, 9), "neq": (15, 8), "div": (307, 109), "rem": (300, 109),
            "overflowing_add": (37, 15), "overflowing_sub": (38, 16)}


def build(op):
    import pzlib
    c = pzlib.Circuit(None)
    a, b, s = c.inputs(8), c.inputs(8), c.inputs(1)
    out = c.u8(op, a, b, s[0] if op == "mux" else None)
    return c, out


@pytest.mark.parametrize("op", sorted(EXPECTED))
def test_gate_count_and_depth(op):
    c, out = build(op)
    st = c.stats()
    assert (st["pbs"], st["depth"]) == EXPECTED[op]
    assert len(out) == {"div_rem": 16, "overflowing_add": 9, "overflowing_sub": 9}.get(
        op, 1 if op in ("lt", "gt", "le", "ge", "eq", "neq") else 8)


def test_second_division_adds_one_or_to_flag():
    import pzlib
    c = pzlib.Circuit(None)
    a, b = c.inputs(8), c.inputs(8)
    c.u8("div", a, b)
    first = c.stats()["pbs"]
    c.u8("div", a, b)
    assert c.stats()["pbs"] - first == 308  # 300 division + 7 is_zero + 1 OR


def test_level_sizes_sum_to_pbs_and_batch_across_ops():
    import pzlib
    c = pzlib.Circuit(None)
    for _ in range(4):
        a, b = c.inputs(8), c.inputs(8)
        c.u8("mul", a, b)
    sizes = c.level_sizes()
    assert int(sizes.sum()) == 4 * 155
    assert len(sizes) == 25 and int(sizes.max()) % 4 == 0
