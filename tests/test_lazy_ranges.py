"""Range claims of the device's lazy Shoup products (phantom-zone_amd/csrc/
pz_arith.hpp), checked in exact integers on the CPU.

shoup_quot4 (twin floor(w 2^64 / q), any a < 2^64) must give a remainder
w*a - q*Q in [0, 4q); shoup_quot6 (twin floor(w 2^63 / q), a < 2^63) a
remainder w*a - 2q*Q in [0, 6q).  The forward transforms FWD_FAST4 / FWD_FAST6
and the key products rely on these bounds; a remainder outside them would
wrap or overflow the next stage's unreduced range.  The functions below follow
the device code's 32-bit partial products one for one."""
import numpy as np

from conftest import rng

Q = 18014398509404161  # 2^54 - 77823
M32 = 0xFFFFFFFF


def quot4(a, ws):
    a0, a1, s0, s1 = a & M32, a >> 32, ws & M32, ws >> 32
    x, y = (s0 * a1) >> 32, (s1 * a0) >> 32
    return (s1 * a1 + x + y) % 2 ** 64


def quot6(a, t):
    a0, a1, t0, t1 = a & M32, a >> 32, t & M32, t >> 32
    col = a1 * t0 + a0 * t1
    assert col < 2 ** 64  # summed in one 64-bit register on the device
    return (a1 * t1 + (col >> 32)) % 2 ** 64


def rem(c, a, w, h, m):
    """c + w*a - m*h mod 2^64 (shoup_rem with modulus m)."""
    return (c + w * a - m * h) % 2 ** 64


def operands(seed, hi):
    r = rng(seed)
    ws = [int(v) for v in r.integers(0, Q, size=400, dtype=np.uint64)] + [0, 1, 2, Q - 1, Q // 2]
    a = [int(v) for v in r.integers(0, hi, size=400, dtype=np.uint64)]
    a += [0, 1, hi - 1, hi // 2, 2 ** 32 - 1, 2 ** 32, 68 * Q - 1, 46 * Q - 1]
    return ws, [v for v in a if v < hi]


def test_quot4_range():
    ws, avs = operands(91, 2 ** 64 - 1)
    for w in ws:
        s = (w << 64) // Q
        for a in avs[::7]:
            r = rem(0, a, w, quot4(a, s), Q)
            assert r < 4 * Q and r % Q == (a * w) % Q


def test_quot6_range():
    ws, avs = operands(92, 2 ** 63)
    for w in ws:
        t = (w << 63) // Q
        assert t >> 32 < 2 ** 31
        for a in avs[::7]:
            r = rem(0, a, w, quot6(a, t), 2 * Q)
            assert r < 6 * Q and r % Q == (a * w) % Q


def test_fast6_forward_bound():
    """FWD_FAST6 butterflies from digits below 2q: x' = x + t, y' = x + 6q - t
    with t < 6q grow the bound by 6q a stage; after the 11 stages of N = 2048
    every value is below 68q < 2^63, inside quot6's operand range."""
    bound = 2 * Q
    for _ in range(11):
        assert bound < 2 ** 63
        bound += 6 * Q
    assert bound == 68 * Q and bound < 2 ** 63


def test_inv6_stage_bounds():
    """The IB = 8 fast inverse of the split step: executed stage k receives
    values below 2^(k-1) 8q and forms dif = x + 2^(k+2) q - y.  Stages 1-6 take
    the 2^63-twin product, so dif must stay below 2^63 there and the [0, 6q)
    product within the next stage's input bound 2^k 8q; stage 7 (KR) may see
    dif up to 2^10 q, beyond the 2^63 operand range, and keeps the 2^64 twin;
    the tail stages 8-10 take the 2^63 twin again."""
    for k in range(1, 7):
        inb = 2 ** (k - 1) * 8 * Q
        dif_max = inb + 2 ** (k + 2) * Q
        assert dif_max <= 2 ** (k + 3) * Q < 2 ** 63
        assert 6 * Q <= 2 ** k * 8 * Q
    assert 2 ** 6 * 8 * Q + 2 ** 9 * Q > 2 ** 63  # stage 7 stays on the 2^64 twin
    # the unreduced tail (stages 8-10 of 11): inputs below 2^(k-6) q, the
    # difference x + 2^(k-6) q - y below 2^(k-5) q; the next stage's bound
    # 2^(k-5) q takes both the sum and the [0, 6q) product
    for k in range(8, 11):
        inb = 2 ** (k - 6) * Q
        assert inb + 2 ** (k - 6) * Q <= 2 ** (k - 5) * Q < 2 ** 63
        assert 2 * inb <= 2 ** (k - 5) * Q and 6 * Q <= 2 ** (k - 5) * Q
    # the product itself at the largest stage-6 difference
    r = rng(93)
    for w in [int(v) for v in r.integers(0, Q, size=300, dtype=np.uint64)] + [1, Q - 1]:
        t = (w << 63) // Q
        for a in (2 ** 9 * Q - 1, 2 ** 9 * Q - 2 ** 40, 2 ** 62):
            res = rem(0, a, w, quot6(a, t), 2 * Q)
            assert res < 6 * Q and res % Q == (a * w) % Q
