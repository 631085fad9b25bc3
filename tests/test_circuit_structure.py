"""FheUint8 circuit structure (no GPU): gate counts and depths of every
operation, built from the C++ restatement of src/shortint/ops.rs.

Expected figures come from tracing the reference's gate calls: mul 155/24,
div_rem 307/109 (including the 7-gate is_zero of the div-by-zero flag), lt
35/9, add 37/15, sub 38/16, eq 15/8, mux 24/2 (PBS / depth)."""
import numpy as np
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


def _check_schedule(c):
    """Every bootstrapped gate runs after the gates producing its inputs (NOT
    chains resolved), every NOT with its input's producer, no level past the
    depth, every level within whole rounds of 512 unless gates whose latest
    level has come force more."""
    nodes = c.nodes()
    lv = c.node_levels()
    st = c.stats()
    src = np.full(len(nodes), -1, dtype=np.int64)
    for i, (kind, op, a, b) in enumerate(nodes):
        if kind == 1:
            src[i] = i
        elif kind == 2:
            src[i] = src[a]
    for i, (kind, op, a, b) in enumerate(nodes):
        if kind == 1:
            for w in (a, b):
                p = src[w]
                assert p < 0 or lv[p] < lv[i], (i, w)
            assert 1 <= lv[i] <= st["depth"]
        elif kind == 2:
            assert lv[i] == (lv[src[i]] if src[i] >= 0 else 0), i
    sizes = c.level_sizes()
    assert int(sizes.sum()) == st["pbs"]
    return sizes


def test_schedule_respects_dependencies_and_fills_levels():
    import pzlib
    c = pzlib.Circuit(None)
    for op, count in (("mul", 64), ("div_rem", 64), ("lt", 64)):
        for _ in range(count):
            a, b = c.inputs(8), c.inputs(8)
            c.u8(op, a, b)
    sizes = _check_schedule(c)
    # as-soon-as-possible levels for comparison: the slack schedule keeps the
    # depth and flattens the early peak into levels of one round (512)
    nodes = c.nodes()
    asap = np.zeros(len(nodes), dtype=np.int64)
    for i, (kind, op, a, b) in enumerate(nodes):
        if kind == 1:
            asap[i] = max(asap[a], asap[b]) + 1
        elif kind == 2:
            asap[i] = asap[a]
    pbs_levels = asap[nodes[:, 0] == 1]
    asap_sizes = np.bincount(pbs_levels)
    assert len(sizes) == len(asap_sizes) == c.stats()["depth"] + 1
    assert int(asap_sizes.max()) > 4 * 512
    assert int((sizes == 512).sum()) >= 40 and int(sizes[1:20].max()) == 512


def test_schedule_of_small_circuits_is_as_soon_as_possible():
    import pzlib
    c = pzlib.Circuit(None)
    a, b = c.inputs(8), c.inputs(8)
    c.u8("add", a, c.u8("mul", a, b))
    sizes = _check_schedule(c)
    assert len(sizes) == c.stats()["depth"] + 1
