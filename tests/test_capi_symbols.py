"""The C-ABI library loads and exports every symbol include/pz.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import PKG, ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pz.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pz_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("pz_ctx_create", "pz_ntt_forward", "pz_shoup_matrix_fma", "pz_decompose",
              "pz_rlwe_by_rgsw", "pz_rlwe_auto", "pz_lwe_key_switch", "pz_pbs", "pz_gates",
              "pz_key_upload"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(PKG, "libpz.so"))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    import pzlib
    assert set(declared_symbols()) == set(pzlib.SIGNATURES)


def test_params_named_without_gpu():
    import pzlib
    p = pzlib.params_named("NI_2P")
    assert (p.rlwe_q, p.rlwe_n, p.lwe_n, p.lwe_q, p.br_q) == (18014398509404161, 2048, 520,
                                                             1 << 16, 1 << 12)
    assert (p.lwe_logb, p.lwe_d, p.rlrg_logb, p.rlrg_da, p.rlrg_db) == (1, 13, 17, 1, 1)
    assert (p.auto_logb, p.auto_d, p.g, p.w) == (24, 1, 5, 10)
    p8 = pzlib.params_named("NI_8P")
    assert (p8.lwe_n, p8.lwe_q, p8.lwe_d) == (660, 1 << 17, 14)


def test_params_match_oracle():
    import pzlib
    from oracle import Oracle, PARAM_SETS
    from oracle.oracle import Params, lib
    for name in PARAM_SETS:
        ours = pzlib.params_named(name)
        theirs = Params()
        lib().pzo_params_named(PARAM_SETS[name], ctypes.byref(theirs))
        for f, _ in Params._fields_:
            assert getattr(ours, f) == getattr(theirs, f), (name, f)
