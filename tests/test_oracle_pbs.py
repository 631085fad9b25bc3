"""Pins of the oracle's PBS and gates (src/pbs.rs, src/bool/evaluator.rs).

* Gate truth tables after decryption for every input pair of all six gates
  (NAND/AND/OR/NOR/XOR/XNOR) and NOT, at NI_2P and SP_TEST.
* The dlog map: (+-)g^dlog[a] == a (mod br_q) for every odd a
  (src/bool/evaluator.rs:664-673).
* The blind-rotation schedule: every LWE index gets exactly one RLWE x RGSW
  product, the automorphism count matches the PBS statistics, and an
  automorphism never follows another without a window reason
  (src/pbs.rs:277-413).
* Test vectors: entries are +-Q/8 and the NAND vector has the partition of
  src/bool/evaluator.rs:690-726 after the X -> X^{-g} map.
"""
import numpy as np
import pytest

from conftest import keys_for, rng

PLAIN = {"nand": lambda a, b: 1 - (a & b), "and": lambda a, b: a & b,
         "or": lambda a, b: a | b, "nor": lambda a, b: 1 - (a | b),
         "xor": lambda a, b: a ^ b, "xnor": lambda a, b: 1 - (a ^ b)}


@pytest.mark.parametrize("name", ["SP_TEST", "NI_2P"])
def test_gate_truth_tables(name):
    o, s, z, key = keys_for(name)
    seed = 1000
    for g, f in PLAIN.items():
        for a in (0, 1):
            for b in (0, 1):
                c0 = o.encrypt_bool(s, a, seed)
                c1 = o.encrypt_bool(s, b, seed + 1)
                seed += 2
                out = o.gate(key, g, c0, c1)
                assert o.decrypt_bool(s, out) == f(a, b), (g, a, b)
    for a in (0, 1):
        c = o.encrypt_bool(s, a, seed)
        assert o.decrypt_bool(s, o.not_(c)) == 1 - a


# the reference's other parameter sets (src/bool/parameters.rs:526-596,
# :652-677): NAND (sum + Q/4) and XOR (2(c0 - c1)), the two linear steps
@pytest.mark.parametrize("name", ["I_2P_LB_SR", "I_4P", "I_8P", "NI_4P_LB_SR"])
def test_gate_truth_tables_other_sets(name):
    o, s, z, key = keys_for(name)
    seed = 3000
    for g in ("nand", "xor"):
        for a in (0, 1):
            for b in (0, 1):
                c0 = o.encrypt_bool(s, a, seed)
                c1 = o.encrypt_bool(s, b, seed + 1)
                seed += 2
                out = o.gate(key, g, c0, c1)
                assert o.decrypt_bool(s, out) == PLAIN[g](a, b), (name, g, a, b)


@pytest.mark.parametrize("name", ["NI_2P", "NI_8P", "NI_4P_HB_FR", "SP_TEST", "I_2P_LB_SR",
                                  "I_8P"])
def test_dlog_map(name):
    from conftest import oracle_for
    o = oracle_for(name)
    brq, g = o.p.br_q, o.p.g
    dl = o.dlog_map()
    for a in range(1, brq, 2):
        k = int(dl[a])
        if k < brq // 4:
            assert pow(g, k, brq) == a
        else:
            assert (brq - pow(g, k - brq // 4, brq)) % brq == a


def test_schedule_and_stats():
    o, s, z, key = keys_for("NI_2P")
    r = rng(70)
    lq = o.p.lwe_q
    for trial in range(20):
        ks = r.integers(0, lq, size=o.n + 1, dtype=np.uint64)
        ops = o.br_schedule(ks)
        kinds = ops & 0xFF
        args = (ops >> 8) & 0x7FFFFF
        ext = sorted(int(a) for k, a in zip(kinds, args) if k == 0)
        assert ext == list(range(o.n))
        autos = int(np.count_nonzero(kinds == 1))
        assert 400 < autos < 560
        assert all(0 <= int(a) <= o.p.w for k, a in zip(kinds, args) if k == 1)
        # the first op is never a non-trivial automorphism without a product
        first_ext = int(np.argmax(kinds == 0))
        assert all((int(op) >> 31) == 1 for op in ops[:first_ext])


def test_pbs_stats_match_schedule():
    o, s, z, key = keys_for("SP_TEST")
    r = rng(71)
    for _ in range(4):
        x = r.integers(0, o.Q, size=o.N + 1, dtype=np.uint64)
        out, st = o.pbs(key, "nand", x)
        assert st.ext_products == o.n
        assert st.autos >= 1


@pytest.mark.parametrize("name", ["NI_2P", "SP_TEST"])
def test_test_vectors(name):
    from conftest import oracle_for
    o = oracle_for(name)
    q = o.Q
    t, f = q >> 3, q - (q >> 3)
    for g in ("nand", "and", "or", "nor", "xor", "xnor"):
        tv = o.test_vec(g)
        assert set(int(v) for v in tv) <= {t, f, q - t, q - f}
    # NAND: before the X -> X^{-g} map, entries < 3*br_q/8 are true
    qby2 = o.p.br_q // 2
    tv = o.test_vec("nand")
    kk = (2 * qby2 - o.p.g) % (2 * qby2)
    for i in range(0, qby2, 5):
        to = (i * kk) % (2 * qby2)
        val = t if i < 3 * (o.p.br_q // 8) else f
        got = int(tv[to]) if to < qby2 else (q - int(tv[to - qby2])) % q
        assert got == val


def py_br_schedule(ks, p):
    """Independent restatement of the blind-rotation op order (src/pbs.rs:
    146-161 buckets, :240-414 loop; dlog map src/bool/evaluator.rs:664-673;
    mod_switch_odd :416-420 in Python floats).  Returns (kind, arg, trivial)
    with kind 0 = RLWE x RGSW of LWE index arg, 1 = automorphism key arg."""
    import math
    br_q, lwe_q, g, w = int(p.br_q), int(p.lwe_q), int(p.g), int(p.w)
    qb4 = br_q // 4
    dlog = [0] * br_q
    for i in range(qb4):
        v = pow(g, i, br_q)
        dlog[v] = i
        dlog[br_q - v] = i + qb4
    buckets = [[] for _ in range(br_q // 2)]
    for index, v in enumerate(int(x) for x in ks[1:]):
        odd = math.floor((float(v) * float(br_q)) / float(lwe_q))
        odd += (odd & 1) ^ 1
        buckets[dlog[odd]].append(index)
    ops, state = [], {"trivial": True}

    def ext(i):
        ops.append((0, i, state["trivial"]))
        state["trivial"] = False

    def auto(k):
        ops.append((1, k, state["trivial"]))

    for base in (qb4, 0):  # -(g^k) buckets, then +(g^k)
        v = 0
        for i in range(qb4 - 1, 0, -1):
            for si in buckets[base + i]:
                ext(si)
            v += 1
            if buckets[base + i - 1] or v == w or i == 1:
                auto(v)
                v = 0
        for si in buckets[base]:
            ext(si)
        if base == qb4:
            auto(0)  # -(g^0) closes with the -g automorphism
    return ops


@pytest.mark.parametrize("name", ["NI_2P", "NI_4P_HB_FR", "NI_8P", "SP_TEST", "I_2P_LB_SR",
                                  "I_4P", "I_8P", "NI_4P_LB_SR"])
def test_schedule_matches_python_restatement(name):
    """The oracle's recorded op order equals an independent restatement of
    the reference's loop, including the trivial-RLWE flags, for uniform key
    switch outputs and for ones concentrated in few buckets."""
    from conftest import oracle_for
    o = oracle_for(name)
    r = rng(72)
    lq = int(o.p.lwe_q)
    cases = [r.integers(0, lq, size=o.n + 1, dtype=np.uint64) for _ in range(6)]
    cases.append(np.full(o.n + 1, 1, dtype=np.uint64))             # one bucket
    cases.append(np.zeros(o.n + 1, dtype=np.uint64))               # +g^0 only
    cases.append(r.integers(0, 4, size=o.n + 1, dtype=np.uint64) * (lq // 4))
    for ks in cases:
        ops = o.br_schedule(ks)
        got = [(int(x) & 0xFF, (int(x) >> 8) & 0x7FFFFF, bool(int(x) >> 31)) for x in ops]
        assert got == py_br_schedule(ks, o.p)


def closed_form_schedule(kbin, nb, w):
    """The op list as the blind-rotation kernel builds it in parallel
    (br_schedule in pz_rlwe.hip): bucket sizes; for bucket k with i = k mod
    br_q/4 >= 1 a pending automorphism fires after it when bucket k-1 is
    non-empty, at i == 1, or where d mod w == 0, with d = one past the next
    non-empty bucket at or above k below the half's top bucket, minus k (the
    half's top + 1 - k if none), and carries v = (d-1) mod w + 1; the first
    half's bucket br_q/4 ends with the v = 0 automorphism.  Each bucket's op
    count then places every op by one exclusive scan (k descending), and the
    indices of a bucket keep ascending order.  Returns (kind, arg) pairs."""
    qb4 = nb // 2
    cnt = [0] * nb
    for k in kbin:
        cnt[k] += 1

    def auto_after(k):
        base = qb4 if k >= qb4 else 0
        i = k - base
        if i == 0:
            return 1 if base else 0
        nat = i == 1 or cnt[k - 1] != 0
        top = base + qb4 - 1
        d = top + 1 - k
        for kk in range(k, top):
            if cnt[kk]:
                d = kk + 1 - k
                break
        return 2 + (d - 1) % w if nat or d % w == 0 else 0

    pos, bpos, slots = 0, [0] * nb, {}
    for k in range(nb - 1, -1, -1):
        a = auto_after(k)
        bpos[k] = pos
        pos += cnt[k]
        if a:
            slots[pos] = (1, a - 1)
            pos += 1
    for i, k in enumerate(kbin):
        rank = sum(1 for j in range(i) if kbin[j] == k)
        slots[bpos[k] + rank] = (0, i)
    assert sorted(slots) == list(range(pos))
    return [slots[x] for x in range(pos)]


@pytest.mark.parametrize("name", ["NI_2P", "NI_4P_HB_FR", "NI_8P", "SP_TEST"])
def test_parallel_schedule_closed_form(name):
    """The closed form the GPU kernel uses to build the op list with the whole
    work group equals the oracle's sequential op order, for uniform key
    switch outputs, every index in one bucket, and sparse fillings with long
    runs of empty buckets (forced automorphisms)."""
    import math
    from conftest import oracle_for
    o = oracle_for(name)
    p = o.p
    br_q, lq = int(p.br_q), int(p.lwe_q)
    qb4, g = br_q // 4, int(p.g)
    dlog = [0] * br_q
    for i in range(qb4):
        v = pow(g, i, br_q)
        dlog[v] = i
        dlog[br_q - v] = i + qb4
    r = rng(73)
    cases = [r.integers(0, lq, size=o.n + 1, dtype=np.uint64) for _ in range(4)]
    cases.append(np.zeros(o.n + 1, dtype=np.uint64))
    cases.append(np.full(o.n + 1, lq // 3, dtype=np.uint64))
    sparse = np.zeros(o.n + 1, dtype=np.uint64)
    sparse[1::37] = r.integers(0, lq, size=len(sparse[1::37]), dtype=np.uint64)
    cases.append(sparse)
    cases.append(r.integers(0, 4, size=o.n + 1, dtype=np.uint64) * (lq // 4) + 1)
    for ks in cases:
        kbin = []
        for v in (int(x) for x in ks[1:]):
            odd = math.floor((float(v) * float(br_q)) / float(lq))
            odd += (odd & 1) ^ 1
            kbin.append(dlog[odd])
        want = [(int(x) & 0xFF, (int(x) >> 8) & 0x7FFFFF) for x in o.br_schedule(ks)]
        assert closed_form_schedule(kbin, br_q // 2, int(p.w)) == want
