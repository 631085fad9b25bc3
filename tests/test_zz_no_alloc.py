"""GPU: pz_gates_noalloc (the gate batch whose pipeline scratch is a fixed
device array) through the C ABI gives the same words as pz_gates on two
calls in a row, the second smaller than the first (so stale scratch from the
first call would show), at NI_2P and at NI_8P's largest batch; outputs
decrypt to the gates; a batch beyond the fixed scratch is refused with its
limit in the message."""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import keys_for

pytestmark = pytest.mark.gpu

PLAIN = {0: lambda a, b: 1 - (a & b), 1: lambda a, b: a & b, 2: lambda a, b: a | b,
         3: lambda a, b: 1 - (a | b), 4: lambda a, b: a ^ b, 5: lambda a, b: 1 - (a ^ b)}


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def batch(ctx, s, B, seed):
    r = np.random.default_rng(seed)
    b0 = r.integers(0, 2, size=B).astype(np.uint8)
    b1 = r.integers(0, 2, size=B).astype(np.uint8)
    ops = r.integers(0, 6, size=B).astype(np.uint32)
    c0 = ctx.to_device(ctx.encrypt_bools(s, b0, seed * 7 + 1))
    c1 = ctx.to_device(ctx.encrypt_bools(s, b1, seed * 7 + 2))
    want = np.array([PLAIN[int(o)](int(a), int(b)) for o, a, b in zip(ops, b0, b1)],
                    dtype=np.uint8)
    return ctx.to_device(ops), c0, c1, want


@pytest.mark.parametrize("name,sizes", [("NI_2P", (300, 37)), ("NI_8P", (512, 200))])
def test_noalloc_gates_equal_pz_gates(name, sizes):
    import pzlib
    o, s, z, key = keys_for(name)
    ctx = pzlib.Context(name)
    try:
        k = ctx.key_upload(key.rgsw, key.auto_key, key.lwe_ksk)
        for i, B in enumerate(sizes):
            log("noalloc %s call %d B=%d" % (name, i, B))
            dops, d0, d1, want = batch(ctx, s, B, 60 + i)
            ref, got = ctx.array((B, ctx.N + 1)), ctx.array((B, ctx.N + 1))
            ctx.gates(k, dops, d0, d1, ref, B)
            rc = ctx.L.pz_gates_noalloc(ctx.h, k.h, dops.ptr, d0.ptr, d1.ptr, got.ptr, B, None)
            assert rc == 0, ctx.L.pz_last_error().decode()
            g = got.download()
            assert np.array_equal(g, ref.download())
            assert np.array_equal(ctx.decrypt_bools(s, g), want)
        big = 577 if name == "NI_2P" else 513
        d = ctx.array((big, ctx.N + 1))
        ops = ctx.to_device(np.zeros(big, np.uint32))
        rc = ctx.L.pz_gates_noalloc(ctx.h, k.h, ops.ptr, d.ptr, d.ptr, d.ptr, big, None)
        assert rc == -1
        assert "at most %d" % (big - 1) in ctx.L.pz_last_error().decode()
        k.free()
    finally:
        ctx.close()
