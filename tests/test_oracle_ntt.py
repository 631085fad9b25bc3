"""Pins of the oracle's NTT (src/ntt.rs) independent of its own code.

* Tables: psi_bo[bitrev(i)] == psi^i, Shoup twins exact, n_inv * N == 1.
* Closed form: forward(a)[j] == a(psi^(2*bitrev(j)+1)) (mod q) -- the
  evaluation of a at the odd powers of psi in bit-reversed order -- checked
  at sampled j for N=2048 and at every j for small rings.
* Products: backward(forward(a) . forward(b)) == the negacyclic product,
  computed here with Python big-integer Kronecker substitution (N = 2048) and
  by schoolbook (the reference's own test, N = 16, 25/40/50/60-bit primes).
* Round trips and output ranges (src/ntt.rs:475-508).
* Lazy values: an independent Python radix-2 restatement of the reference's
  loops over Python ints, for every variant at N = 64.
"""
import numpy as np
import pytest

from conftest import oracle_for, rng
from oracle.oracle import lib


def bitrev(i, logn):
    return int(format(i, "0%db" % logn)[::-1], 2)


def tables(q, n):
    L = lib()
    t = L.pzo_ntt_new(q, n)
    out = np.zeros(4 * n + 2, dtype=np.uint64)
    L.pzo_ntt_tables(t, out)
    psi = L.pzo_ntt_psi(t)
    return t, psi, out


def run(t, which, x):
    a = np.ascontiguousarray(x, dtype=np.uint64).copy()
    n = a.shape[-1]
    lib().pzo_ntt_rows(t, which, a.reshape(-1), a.size // n)
    return a


def find_prime(bits, modulo):
    """Largest prime below 2^bits that is 1 mod `modulo` (generate_prime,
    src/utils.rs:122-142), via Miller-Rabin with fixed bases."""
    def is_prime(p):
        if p < 2:
            return False
        for sp in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            if p % sp == 0:
                return p == sp
        d, s = p - 1, 0
        while d % 2 == 0:
            d //= 2
            s += 1
        for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            x = pow(a, d, p)
            if x in (1, p - 1):
                continue
            for _ in range(s - 1):
                x = x * x % p
                if x == p - 1:
                    break
            else:
                return False
        return True
    p = (1 << bits) - 1
    while p % modulo != 1:
        p -= 1
    while not is_prime(p):
        p -= modulo
    return p


@pytest.mark.parametrize("name", ["NI_2P", "SP_TEST"])
def test_tables(name):
    o = oracle_for(name)
    q, n = o.Q, o.N
    tab = o.ntt_tables()
    psi = o.psi()
    logn = n.bit_length() - 1
    psi_inv = pow(psi, q - 2, q)
    for i in range(n):
        bo = bitrev(i, logn)
        assert int(tab[bo]) == pow(psi, i, q)
        assert int(tab[2 * n + bo]) == pow(psi_inv, i, q)
    for k in range(4 * n):
        if k // n in (1, 3):
            assert int(tab[k]) == (int(tab[k - n]) << 64) // q
    assert int(tab[4 * n]) * n % q == 1
    assert int(tab[4 * n + 1]) == (int(tab[4 * n]) << 64) // q


@pytest.mark.parametrize("name", ["NI_2P", "SP_TEST"])
def test_forward_closed_form(name):
    o = oracle_for(name)
    q, n = o.Q, o.N
    psi = o.psi()
    logn = n.bit_length() - 1
    r = rng(7)
    a = r.integers(0, q, size=n, dtype=np.uint64)
    ai = [int(v) for v in a]
    fwd = o.ntt_rows("forward", a)
    lazy = o.ntt_rows("forward_lazy", a)
    assert int(fwd.max()) < q and int(lazy.max()) < 2 * q
    sample = range(n) if n <= 512 else list(range(0, n, 97)) + [n - 1]
    for j in sample:
        root = pow(psi, 2 * bitrev(j, logn) + 1, q)
        ev = 0
        for c in reversed(ai):  # Horner
            ev = (ev * root + c) % q
        assert int(fwd[j]) == ev
        assert int(lazy[j]) % q == ev


def negacyclic_bigint(a, b, q):
    """Negacyclic product over Python ints (Kronecker substitution)."""
    n = len(a)
    slot = 2 * q.bit_length() + n.bit_length() + 2
    A = sum(int(v) << (slot * i) for i, v in enumerate(a))
    B = sum(int(v) << (slot * i) for i, v in enumerate(b))
    P = A * B
    mask = (1 << slot) - 1
    coef = [(P >> (slot * i)) & mask for i in range(2 * n - 1)] + [0]
    return [(coef[i] - coef[i + n]) % q for i in range(n)]


@pytest.mark.parametrize("name", ["NI_2P", "SP_TEST"])
def test_product_matches_bigint_negacyclic(name):
    o = oracle_for(name)
    q, n = o.Q, o.N
    L = lib()
    r = rng(8)
    for _ in range(2):
        a = r.integers(0, q, size=n, dtype=np.uint64)
        b = r.integers(0, q, size=n, dtype=np.uint64)
        fa = o.ntt_rows("forward_lazy", a)
        fb = o.ntt_rows("forward_lazy", b)
        prod = np.array([L.pzo_mul_mod_fast(q, int(x), int(y)) for x, y in zip(fa, fb)],
                        dtype=np.uint64)
        back = o.ntt_rows("backward", prod)
        assert [int(v) for v in back] == negacyclic_bigint(a, b, q)


def test_reference_negacyclic_small_primes():
    """src/ntt.rs:511-540: N=16, primes of 25/40/50/60 bits."""
    L = lib()
    n = 16
    r = rng(9)
    for bits in (25, 40, 50, 60):
        q = find_prime(bits, 2 * n)
        t, psi, _ = tables(q, n)
        assert pow(psi, n, q) == q - 1
        for _ in range(16):
            a = r.integers(0, q, size=n, dtype=np.uint64)
            b = r.integers(0, q, size=n, dtype=np.uint64)
            fa, fb = run(t, 0, a), run(t, 0, b)
            prod = np.array([L.pzo_mul_mod_fast(q, int(x), int(y)) for x, y in zip(fa, fb)],
                            dtype=np.uint64)
            back = run(t, 3, prod)
            school = np.zeros(n, dtype=np.uint64)
            L.pzo_negacyclic_mul_schoolbook(q, a, b, school, n)
            assert np.array_equal(back, school)
            assert [int(v) for v in back] == negacyclic_bigint(a, b, q)
        L.pzo_ntt_free(t)


@pytest.mark.parametrize("name", ["NI_2P", "SP_TEST"])
def test_round_trips_and_ranges(name):
    """src/ntt.rs:475-508."""
    o = oracle_for(name)
    q = o.Q
    r = rng(10)
    a = r.integers(0, q, size=(8, o.N), dtype=np.uint64)
    f = o.ntt_rows("forward", a)
    assert int(f.max()) < q and not np.array_equal(f, a)
    assert np.array_equal(o.ntt_rows("backward", f), a)
    fl = o.ntt_rows("forward_lazy", a)
    assert int(fl.max()) < 2 * q
    assert np.array_equal(o.ntt_rows("backward", fl), a)
    bl = o.ntt_rows("backward_lazy", f)
    assert int(bl.max()) < 2 * q
    assert np.array_equal(bl % np.uint64(q), a)


def py_ntt(a, psi, psis, q, strict):
    """Python restatement of ntt_lazy / ntt (src/ntt.rs:110-180)."""
    a = [int(v) for v in a]
    n = len(a)
    q2 = 2 * q
    M64 = (1 << 64) - 1

    def smul(y, w, ws):
        return (w * y - q * ((ws * y) >> 64)) & M64

    t, m = n, 1
    while m < n:
        t >>= 1
        for i in range(m):
            w, ws = int(psi[m + i]), int(psis[m + i])
            for j in range(t):
                x, y = a[2 * i * t + j], a[2 * i * t + j + t]
                if x >= q2:
                    x -= q2
                tt = smul(y, w, ws)
                if t == 1:
                    ox, oy = (x + tt) & M64, (x - tt) & M64
                    ox = min(ox, (ox - q2) & M64)
                    oy = min(oy, (oy + q2) & M64)
                    if strict:
                        ox = min(ox, (ox - q) & M64)
                        oy = min(oy, (oy - q) & M64)
                else:
                    ox, oy = x + tt, x + q2 - tt
                a[2 * i * t + j], a[2 * i * t + j + t] = ox, oy
        m <<= 1
    return a


def py_intt(a, ipsi, ipsis, ninv, ninvs, q, strict):
    """Python restatement of ntt_inv_lazy / ntt_inv (src/ntt.rs:187-278)."""
    a = [int(v) for v in a]
    n = len(a)
    q2 = 2 * q
    M64 = (1 << 64) - 1

    def smul(y, w, ws):
        return (w * y - q * ((ws * y) >> 64)) & M64

    m, t = n >> 1, 1
    while m > 0:
        for i in range(m):
            w, ws = int(ipsi[m + i]), int(ipsis[m + i])
            for j in range(t):
                x, y = a[2 * i * t + j], a[2 * i * t + j + t]
                xd = x + y
                if xd >= q2:
                    xd -= q2
                yd = smul(x + q2 - y, w, ws)
                if m == 1:
                    xd, yd = smul(xd, ninv, ninvs), smul(yd, ninv, ninvs)
                    if strict:
                        xd, yd = min(xd, (xd - q) & M64), min(yd, (yd - q) & M64)
                a[2 * i * t + j], a[2 * i * t + j + t] = xd, yd
        t *= 2
        m >>= 1
    return a


def test_lazy_values_match_python_restatement():
    q = find_prime(54, 128)
    n = 64
    t, psi, tab = tables(q, n)
    r = rng(11)
    for _ in range(8):
        a = r.integers(0, q, size=n, dtype=np.uint64)
        a2 = r.integers(0, 2 * q, size=n, dtype=np.uint64)
        assert [int(v) for v in run(t, 0, a)] == py_ntt(a, tab[:n], tab[n:2 * n], q, False)
        assert [int(v) for v in run(t, 1, a)] == py_ntt(a, tab[:n], tab[n:2 * n], q, True)
        ninv, ninvs = int(tab[4 * n]), int(tab[4 * n + 1])
        assert [int(v) for v in run(t, 2, a2)] == py_intt(a2, tab[2 * n:3 * n], tab[3 * n:4 * n],
                                                         ninv, ninvs, q, False)
        assert [int(v) for v in run(t, 3, a2)] == py_intt(a2, tab[2 * n:3 * n], tab[3 * n:4 * n],
                                                         ninv, ninvs, q, True)
    lib().pzo_ntt_free(t)
