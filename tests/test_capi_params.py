"""Parameter validation of pz_ctx_create, through the C ABI and without a
GPU: every rejected set fails before the library touches a device, with the
status the header documents (PZ_ERR_ARG for out-of-range values,
PZ_ERR_UNSUPPORTED for shapes the kernels are not built for)."""
import ctypes

import pytest

PZ_ERR_ARG, PZ_ERR_UNSUPPORTED = -1, -3


def create(p):
    """pz_ctx_create's status and error text (the text tells a rejected
    parameter from a missing device, which is also PZ_ERR_ARG)."""
    import pzlib
    L = pzlib.load()
    h = ctypes.c_void_p()
    rc = L.pz_ctx_create(ctypes.byref(p), 0, ctypes.byref(h))
    err = L.pz_last_error() or b""
    return rc, err


def named(name="NI_2P"):
    import pzlib
    return pzlib.params_named(name)


CASES = [
    ("rlwe_q", lambda p: setattr(p, "rlwe_q", (1 << 54) + 1), PZ_ERR_ARG),       # above 2^54
    ("rlwe_q", lambda p: setattr(p, "rlwe_q", p.rlwe_q + 2), PZ_ERR_ARG),         # not 1 mod 2N
    ("rlwe_q", lambda p: setattr(p, "rlwe_q", 4096 * 1000 + 1), PZ_ERR_UNSUPPORTED),  # < 2^32 at N = 2048
    ("rlwe_n", lambda p: setattr(p, "rlwe_n", 1024), PZ_ERR_UNSUPPORTED),
    ("rlwe_n", lambda p: setattr(p, "rlwe_n", 3000), PZ_ERR_UNSUPPORTED),
    ("lwe_q", lambda p: setattr(p, "lwe_q", 3 << 14), PZ_ERR_UNSUPPORTED),
    ("lwe_q", lambda p: setattr(p, "lwe_q", 1 << 18), PZ_ERR_UNSUPPORTED),
    ("br_q", lambda p: setattr(p, "br_q", 2 * p.rlwe_n * 2), PZ_ERR_ARG),
    ("br_q", lambda p: setattr(p, "br_q", 3000), PZ_ERR_ARG),
    ("w", lambda p: setattr(p, "w", 200), PZ_ERR_UNSUPPORTED),
    ("lwe_n", lambda p: setattr(p, "lwe_n", 0), PZ_ERR_UNSUPPORTED),
    ("lwe_logb", lambda p: setattr(p, "lwe_logb", 8), PZ_ERR_UNSUPPORTED),
    ("rlrg_da", lambda p: setattr(p, "rlrg_da", 0), PZ_ERR_ARG),
    ("rlrg_logb", lambda p: setattr(p, "rlrg_logb", 31), PZ_ERR_UNSUPPORTED),     # base above 2^30
    ("auto_logb", lambda p: setattr(p, "auto_logb", 0), PZ_ERR_UNSUPPORTED),
    ("auto_d", lambda p: setattr(p, "auto_d", 3), PZ_ERR_ARG),                    # 24*3 > 54
    ("lwe_d", lambda p: setattr(p, "lwe_d", 17), PZ_ERR_ARG),                     # 1*17 > 16
    ("uitos_d", lambda p: setattr(p, "uitos_d", 55), PZ_ERR_ARG),
    ("rgrg_da", lambda p: setattr(p, "rgrg_da", 14), PZ_ERR_ARG),                 # 4*14 > 54
    ("rgrg_db", lambda p: setattr(p, "rgrg_db", 0), PZ_ERR_ARG),                  # one count only
]


@pytest.mark.parametrize("field,mutate,status", CASES)
def test_invalid_params_rejected_without_gpu(field, mutate, status):
    p = named()
    mutate(p)
    rc, err = create(p)
    assert rc == status, (field, err)
    assert b"device" not in err, (field, err)


def test_null_arguments():
    import pzlib
    L = pzlib.load()
    h = ctypes.c_void_p()
    assert L.pz_ctx_create(None, 0, ctypes.byref(h)) == PZ_ERR_ARG
    p = named()
    assert L.pz_ctx_create(ctypes.byref(p), 0, None) == PZ_ERR_ARG
    assert L.pz_params_named(None, ctypes.byref(p)) == PZ_ERR_ARG
    assert L.pz_params_named(b"NO_SUCH_SET", ctypes.byref(p)) != 0
