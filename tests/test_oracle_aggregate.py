"""Oracle pin for rgsw_by_rgsw_inplace (src/rgsw/runtime.rs:883-985), the
product the non-interactive server-key aggregation runs per LWE index and
party (src/bool/evaluator.rs:1505-1535).

Closed form over Python ints: for every RLWE row pair (a, b) of RGSW0 (rows
(i, d_a + i) of the -sm part, then (2d_a + i, 2d_a + d_b + i) of the m part),
with digits of a and b under the RGSW x RGSW decomposer and the
coefficient-domain rows of RGSW1,
  a' = sum_r d_r(a) * A(-sm)_r + sum_r d_r(b) * A(m)_r,
  b' = sum_r d_r(a) * B(-sm)_r + sum_r d_r(b) * B(m)_r   in Z_q[X]/(X^N + 1);
the outputs are canonical, so this fixes every word."""
import pytest

from conftest import oracle_for, rng, uniform_mod
from test_oracle_decomposer import py_digits
from test_oracle_ntt import negacyclic_bigint


def _rows_sum(digit_rows, key_rows, q):
    acc = [0] * len(key_rows[0])
    for dr, kr in zip(digit_rows, key_rows):
        prod = negacyclic_bigint(dr, [int(v) for v in kr], q)
        acc = [(x + y) % q for x, y in zip(acc, prod)]
    return acc


@pytest.mark.parametrize("name", ["NI_2P", "I_2P_LB_SR"])
def test_rgsw_by_rgsw_closed_form(name):
    o = oracle_for(name)
    p, q, N = o.p, o.Q, o.N
    da0, db0, da1, db1, lb = p.rlrg_da, p.rlrg_db, p.rgrg_da, p.rgrg_db, p.rgrg_logb
    r = rng(320)
    rgsw0 = uniform_mod(r, q, (2 * da0 + 2 * db0, N))
    rgsw0[0, :3] = [0, q - 1, q // 2]
    rgsw1 = uniform_mod(r, q, (2 * da1 + 2 * db1, N))
    got = o.rgsw_by_rgsw(rgsw0, rgsw1)
    k1 = o.ntt_rows("backward", rgsw1)
    nsm_a, nsm_b = k1[:da1], k1[da1:2 * da1]
    m_a, m_b = k1[2 * da1:2 * da1 + db1], k1[2 * da1 + db1:]
    pairs = [(i, da0 + i) for i in range(da0)] + \
            [(2 * da0 + i, 2 * da0 + db0 + i) for i in range(db0)]
    for ra, rb in pairs:
        dig_a = [py_digits(q, lb, da1, int(v)) for v in rgsw0[ra]]
        dig_b = [py_digits(q, lb, db1, int(v)) for v in rgsw0[rb]]
        rows_a = [[dg[k] % q for dg in dig_a] for k in range(da1)]
        rows_b = [[dg[k] % q for dg in dig_b] for k in range(db1)]
        want_a = [(x + y) % q for x, y in zip(_rows_sum(rows_a, nsm_a, q), _rows_sum(rows_b, m_a, q))]
        want_b = [(x + y) % q for x, y in zip(_rows_sum(rows_a, nsm_b, q), _rows_sum(rows_b, m_b, q))]
        assert got[ra].tolist() == want_a
        assert got[rb].tolist() == want_b
