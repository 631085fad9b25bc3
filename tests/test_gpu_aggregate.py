"""GPU RGSW x RGSW product of the non-interactive server-key aggregation
(rgsw_by_rgsw_inplace, src/rgsw/runtime.rs:883-985) against the oracle, word
for word: uniform coefficient-domain RGSW0 and evaluation-domain RGSW1 under
every parameter set with an RGSW x RGSW decomposer, a batch of one pair per
LWE index of NI_2P (the shape of one party's aggregation step) checked on a
sample, and the error for a parameter set without that decomposer."""
import sys

import numpy as np
import pytest

from conftest import oracle_for, rng, uniform_mod

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pairs(o, r, B):
    p, q, N = o.p, o.Q, o.N
    r0 = 2 * (p.rlrg_da + p.rlrg_db)
    r1 = 2 * (p.rgrg_da + p.rgrg_db)
    g0 = uniform_mod(r, q, (B, r0, N))
    g1 = uniform_mod(r, q, (B, r1, N))
    g0[0, 0, :4] = [0, q - 1, q // 2, q // 2 + 1]  # decomposer edge values
    return g0, g1


@pytest.mark.parametrize("name", ["NI_2P", "NI_4P_HB_FR", "NI_8P", "NI_4P_LB_SR", "I_2P_LB_SR",
                                  "I_4P", "I_8P"])
def test_rgsw_by_rgsw_matches_oracle(name):
    import pzlib
    o = oracle_for(name)
    B = 3
    log("rgsw x rgsw %s B=%d" % (name, B))
    g0, g1 = _pairs(o, rng(330), B)
    ctx = pzlib.Context(name)
    try:
        d0, d1 = ctx.to_device(g0), ctx.to_device(g1)
        ctx.rgsw_by_rgsw(d0, d1, B)
        got = d0.download()
        assert np.array_equal(d1.download(), g1)  # RGSW1 is read-only
    finally:
        ctx.close()
    for b in range(B):
        assert np.array_equal(got[b], o.rgsw_by_rgsw(g0[b], g1[b])), b


def test_rgsw_by_rgsw_lwe_batch_ni2p():
    import pzlib
    o = oracle_for("NI_2P")
    B = o.p.lwe_n
    log("rgsw x rgsw NI_2P B=%d" % B)
    g0, g1 = _pairs(o, rng(331), B)
    ctx = pzlib.Context("NI_2P")
    try:
        d0, d1 = ctx.to_device(g0), ctx.to_device(g1)
        ctx.rgsw_by_rgsw(d0, d1, B)
        got = d0.download()
    finally:
        ctx.close()
    for b in [0, 1, B // 2, B - 1]:
        assert np.array_equal(got[b], o.rgsw_by_rgsw(g0[b], g1[b])), b


def test_rgsw_by_rgsw_unsupported():
    import pzlib
    o = oracle_for("SP_TEST")
    ctx = pzlib.Context("SP_TEST")
    try:
        r0 = 2 * (o.p.rlrg_da + o.p.rlrg_db)
        d0 = ctx.array((1, r0, o.N))
        with pytest.raises(pzlib.PzError):
            ctx.rgsw_by_rgsw(d0, d0, 1)
    finally:
        ctx.close()
