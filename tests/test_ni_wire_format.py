"""CPU: the seeded wire formats of the non-interactive server key (no device
needed).  The blob sizes and the header layout equal what the format states
in include/pz.h, derived here from the parameter set: a 272-byte header,
then the share's seven sections (leader RGSW rows with the RLWE x RGSW
count, the other indices with the RGSW x RGSW count, zero encryptions, u->s
b parts, zero encryptions for the others, automorphism-key and LWE-KSK b
parts) or the seeded key's four; sets without the non-interactive
decomposers are refused."""
import ctypes as C

import pytest

HEADER = 272


def share_words(p, me, total):
    N, n = p.rlwe_n, p.lwe_n
    per = -(-n // total)
    lead = max(0, min(per * (me + 1), n) - min(per * me, n))
    d1, dm = max(p.rlrg_da, p.rlrg_db), max(p.rgrg_da, p.rgrg_db)
    return [lead * d1 * N, (n - lead) * dm * N, n * p.rgrg_da * N, p.uitos_d * N,
            (total - 1) * p.uitos_d * N, (p.w + 1) * p.auto_d * N, N * p.lwe_d]


def key_words(p, total):
    N = p.rlwe_n
    return [total * p.uitos_d * N, p.lwe_n * (2 * p.rlrg_da + 2 * p.rlrg_db) * N,
            (p.w + 1) * p.auto_d * N, N * p.lwe_d]


@pytest.mark.parametrize("name,total", [("NI_2P", 2), ("NI_4P_HB_FR", 3), ("NI_4P_LB_SR", 4),
                                        ("NI_8P", 8), ("NI_8P", 5)])
def test_blob_sizes(name, total):
    import pzlib
    p = pzlib.params_named(name)
    for me in range(total):
        assert pzlib.ni_share_bytes(p, me, total) == HEADER + 8 * sum(share_words(p, me, total))
    assert pzlib.ni_seeded_key_bytes(p, total) == HEADER + 8 * sum(key_words(p, total))


def test_header_is_packed():
    import pzlib
    assert C.sizeof(pzlib.Params) == 88
    assert C.sizeof(pzlib.NiBlobHeader) == HEADER


def test_refusals():
    import pzlib
    sp = pzlib.params_named("SP_TEST")
    with pytest.raises(pzlib.PzError):
        pzlib.ni_share_bytes(sp, 0, 2)
    with pytest.raises(pzlib.PzError):
        pzlib.ni_seeded_key_bytes(sp, 2)
    ni = pzlib.params_named("NI_2P")
    with pytest.raises(pzlib.PzError):
        pzlib.ni_share_bytes(ni, 2, 2)  # user id out of range
    with pytest.raises(pzlib.PzError):
        pzlib.ni_seeded_key_bytes(ni, 9)  # more than PZ_NI_MAX_PARTIES
