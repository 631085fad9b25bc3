"""CPU: every entry point of include/pz.h that rejects its arguments leaves a
message for pz_last_error() naming the function that refused them.  Each
entry point is called with null handles and buffers and zero sizes (no
device is touched: argument checks come first); the message must be
non-empty and new (a different failing call sets a known message before each
one).  Entry points documented to accept NULL (destroy / free) return PZ_OK;
pz_circuit_create with a null context makes a build-only circuit."""
import ctypes as C

import pytest

NULL_OK = {"pz_ctx_destroy", "pz_key_free", "pz_circuit_destroy", "pz_server_destroy",
           "pz_rlwe_ksk_free", "pz_last_error"}


def null_args(argtypes):
    out = []
    for t in argtypes:
        if t in (C.c_int, C.c_uint32, C.c_size_t, C.c_uint64):
            out.append(0)
        elif t is C.c_double:
            out.append(0.0)
        else:
            out.append(None)
    return out


def entry_points():
    import pzlib
    return sorted(pzlib.SIGNATURES)


@pytest.mark.parametrize("name", entry_points())
def test_argument_errors_set_a_message(name):
    import pzlib
    L = pzlib.load()
    if name in NULL_OK:
        if name != "pz_last_error":
            assert getattr(L, name)(None) == 0
        return
    if name == "pz_circuit_create":
        h = C.c_void_p()
        assert L.pz_circuit_create(None, C.byref(h)) == 0 and h.value
        assert L.pz_circuit_destroy(h) == 0
        assert L.pz_circuit_create(None, None) == -1  # PZ_ERR_ARG
        assert L.pz_last_error().decode().startswith("pz_circuit_create:")
        return
    # a known message first (from another entry point), so that a stale one
    # cannot pass
    if name == "pz_params_named":
        assert L.pz_sync(None) < 0
    else:
        assert L.pz_params_named(b"no such set", None) < 0
    marker = L.pz_last_error().decode()
    assert marker
    argtypes = pzlib.SIGNATURES[name][1]
    rc = getattr(L, name)(*null_args(argtypes))
    assert rc < 0, (name, rc)
    msg = L.pz_last_error().decode()
    assert msg and msg != marker, (name, msg)
    assert ": " in msg, (name, msg)
