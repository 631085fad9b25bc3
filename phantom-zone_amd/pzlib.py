"""ctypes binding of libpz.so (the C ABI declared in include/pz.h).

This is the Python face of the MI355X bootstrapping engine: device buffers,
contexts, resident server keys and the batched operators.  It never falls back
to the CPU: if libpz.so is missing or no GPU is visible, loading fails loudly.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PZ_LIBPZ") or os.path.join(HERE, "libpz.so")

GATES = {"nand": 0, "and": 1, "or": 2, "nor": 3, "xor": 4, "xnor": 5}


class PzError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("rlwe_q", C.c_uint64), ("rlwe_n", C.c_uint32), ("lwe_n", C.c_uint32),
        ("lwe_q", C.c_uint64), ("br_q", C.c_uint32),
        ("lwe_logb", C.c_uint32), ("lwe_d", C.c_uint32),
        ("rlrg_logb", C.c_uint32), ("rlrg_da", C.c_uint32), ("rlrg_db", C.c_uint32),
        ("auto_logb", C.c_uint32), ("auto_d", C.c_uint32),
        ("g", C.c_uint32), ("w", C.c_uint32), ("lwe_secret_gaussian", C.c_uint32),
        ("uitos_logb", C.c_uint32), ("uitos_d", C.c_uint32),
        ("rgrg_logb", C.c_uint32), ("rgrg_da", C.c_uint32), ("rgrg_db", C.c_uint32),
    ]


class NiBlobHeader(C.Structure):
    """pz_ni_blob_header: the header of a seeded share / server key blob."""
    _fields_ = [("magic", C.c_char * 8), ("version", C.c_uint32), ("user", C.c_uint32),
                ("total", C.c_uint32), ("nsections", C.c_uint32), ("params", Params),
                ("crs", C.c_uint8 * 32), ("offset", C.c_uint64 * 8), ("words", C.c_uint64 * 8)]


def ni_blob_sections(blob):
    """(header, [u64 arrays of the sections]) of a wire-format blob."""
    h = NiBlobHeader.from_buffer_copy(bytes(blob[:C.sizeof(NiBlobHeader)]))
    secs = [np.frombuffer(bytes(blob[h.offset[i]:h.offset[i] + 8 * h.words[i]]), dtype=np.uint64)
            for i in range(h.nsections)]
    return h, secs


# Every symbol declared in include/pz.h, with its ctypes signature.
_vp = C.c_void_p
SIGNATURES = {
    "pz_params_named": (C.c_int, [C.c_char_p, C.POINTER(Params)]),
    "pz_last_error": (C.c_char_p, []),
    "pz_ctx_create": (C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(_vp)]),
    "pz_ctx_destroy": (C.c_int, [_vp]),
    "pz_ctx_info": (C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                              C.POINTER(C.c_uint32)]),
    "pz_ctx_ntt_tables": (C.c_int, [_vp, _vp]),
    "pz_ctx_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "pz_malloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "pz_free": (C.c_int, [_vp, _vp]),
    "pz_memcpy_htod": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "pz_memcpy_dtoh": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "pz_sync": (C.c_int, [_vp]),
    "pz_ntt_forward": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "pz_ntt_forward_lazy": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "pz_ntt_backward": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "pz_ntt_backward_lazy": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "pz_shoup_matrix_fma": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_size_t, _vp]),
    "pz_decompose": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_size_t, _vp]),
    "pz_key_upload": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "pz_key_free": (C.c_int, [_vp]),
    "pz_key_bytes": (C.c_int, [_vp, C.POINTER(C.c_size_t)]),
    "pz_key_download": (C.c_int, [_vp, _vp, _vp, _vp]),
    "pz_keygen_secrets": (C.c_int, [_vp, _vp, C.c_uint32, _vp, _vp]),
    "pz_keygen_server": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "pz_encrypt_bools": (C.c_int, [_vp, _vp, C.c_uint32, _vp, _vp, C.c_size_t, _vp]),
    "pz_decrypt_bools": (C.c_int, [_vp, _vp, C.c_uint32, _vp, C.c_size_t, _vp]),
    "pz_circuit_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "pz_circuit_destroy": (C.c_int, [_vp]),
    "pz_circuit_inputs": (C.c_int, [_vp, C.c_uint32, _vp]),
    "pz_circuit_gate": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, _vp]),
    "pz_circuit_u8": (C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp, _vp, _vp]),
    "pz_circuit_div_zero_flag": (C.c_int, [_vp, _vp]),
    "pz_circuit_reset_flags": (C.c_int, [_vp]),
    "pz_circuit_stats": (C.c_int, [_vp, _vp, _vp, _vp]),
    "pz_circuit_level_sizes": (C.c_int, [_vp, _vp, C.c_uint32]),
    "pz_circuit_nodes": (C.c_int, [_vp, _vp, C.c_size_t]),
    "pz_circuit_node_levels": (C.c_int, [_vp, _vp, C.c_size_t]),
    "pz_circuit_run": (C.c_int, [_vp, _vp, _vp, _vp]),
    "pz_circuit_read": (C.c_int, [_vp, _vp, C.c_size_t, _vp, _vp]),
    "pz_server_create": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "pz_server_destroy": (C.c_int, [_vp]),
    "pz_server_add_session": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "pz_server_add_session_ctx": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "pz_server_submit": (C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp, C.c_uint32, C.c_uint32,
                                   C.c_double, _vp]),
    "pz_server_run": (C.c_int, [_vp]),
    "pz_server_intake": (C.c_int, [_vp, C.c_int]),
    "pz_server_result": (C.c_int, [_vp, C.c_uint32, _vp]),
    "pz_server_request_times": (C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp, _vp, _vp]),
    "pz_server_stats": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "pz_rlwe_by_rgsw": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "pz_rlwe_auto": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "pz_lwe_key_switch": (C.c_int, [_vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "pz_pbs": (C.c_int, [_vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "pz_gates": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "pz_gates_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "pz_not": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "pz_last_auto_counts": (C.c_int, [_vp, _vp, C.c_size_t]),
    "pz_set_timing": (C.c_int, [_vp, C.c_int]),
    "pz_last_timing": (C.c_int, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.POINTER(C.c_float)]),
    "pz_valu_probe": (C.c_int, [_vp, C.c_int, C.c_uint32, C.POINTER(C.c_double)]),
    "pz_set_blind_rotate_mode": (C.c_int, [_vp, C.c_int]),
    "pz_keygen_user_secret": (C.c_int, [_vp, _vp, _vp]),
    "pz_keygen_rlwe_ksk": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "pz_rlwe_ksk_upload": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "pz_rlwe_ksk_download": (C.c_int, [_vp, _vp]),
    "pz_rlwe_ksk_free": (C.c_int, [_vp]),
    "pz_encrypt_rlwe_bools": (C.c_int, [_vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "pz_rlwe_key_switch": (C.c_int, [_vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "pz_sample_extract": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp, _vp]),
    "pz_rgsw_by_rgsw": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp]),
    "pz_autotune": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "pz_ni_share_uitos": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_uint32, _vp, _vp, _vp, _vp]),
    "pz_ni_share_rgsw": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_uint32, _vp, _vp, _vp,
                                   C.c_uint32, C.c_uint32, _vp, _vp]),
    "pz_ni_uitos_ksk": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, _vp, _vp,
                                  C.POINTER(_vp)]),
    "pz_ni_aggregate_rgsw": (C.c_int, [_vp, _vp, C.c_uint32, _vp, C.c_uint32, C.c_uint32, _vp,
                                       _vp, _vp, _vp]),
    "pz_ni_share_auto_ksk": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pz_ni_share_bytes": (C.c_int, [C.POINTER(Params), C.c_uint32, C.c_uint32,
                                    C.POINTER(C.c_size_t)]),
    "pz_ni_seeded_key_bytes": (C.c_int, [C.POINTER(Params), C.c_uint32, C.POINTER(C.c_size_t)]),
    "pz_ni_key_share": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_uint32, _vp, _vp, _vp, _vp]),
    "pz_ni_aggregate_shares": (C.c_int, [_vp, _vp, C.c_uint32, _vp]),
    "pz_ni_server_key_from_seeded": (C.c_int, [_vp, _vp, C.POINTER(_vp), _vp]),
    "pz_ni_aggregate_auto_ksk": (C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp, _vp, _vp]),
    "pz_ni_expand_server_key": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(_vp)]),
}

_lib = None


def build(jobs=3):
    """Compile libpz.so with hipcc for gfx950 (phantom-zone_amd/Makefile)."""
    subprocess.run(["make", "-s", "-C", HERE, "-j%d" % jobs], check=True)


def load():
    """Load libpz.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PzError("libpz.so not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise PzError("pz error %d: %s" % (rc, load().pz_last_error().decode()))


def aligned_bytes(nbytes):
    """A zeroed uint8 buffer of nbytes whose data is 8-byte aligned."""
    return np.zeros((nbytes + 7) // 8, dtype=np.uint64).view(np.uint8)[:nbytes]


def ni_share_bytes(params, me, total):
    p = params_named(params) if isinstance(params, str) else params
    n = C.c_size_t()
    _check(load().pz_ni_share_bytes(C.byref(p), me, total, C.byref(n)))
    return n.value


def ni_seeded_key_bytes(params, total):
    p = params_named(params) if isinstance(params, str) else params
    n = C.c_size_t()
    _check(load().pz_ni_seeded_key_bytes(C.byref(p), total, C.byref(n)))
    return n.value


def params_named(name):
    p = Params()
    _check(load().pz_params_named(name.encode(), C.byref(p)))
    return p


class DeviceArray:
    """A device allocation with numpy-shaped host transfers."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        ptr = C.c_void_p()
        _check(ctx.L.pz_malloc(ctx.h, max(self.nbytes, 8), C.byref(ptr)))
        self.ptr = ptr.value

    def upload(self, host):
        a = np.ascontiguousarray(host, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        _check(self.ctx.L.pz_memcpy_htod(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes))
        return self

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(self.ctx.L.pz_memcpy_dtoh(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr is not None:
            self.ctx.L.pz_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


U8_OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "rem": 4, "div_rem": 5, "eq": 6, "neq": 7,
          "lt": 8, "gt": 9, "le": 10, "ge": 11, "mux": 12, "max": 13, "min": 14,
          "overflowing_add": 15, "overflowing_sub": 16}
GATE_NOT = 6


class Circuit:
    """A gate DAG of FheUint8 operations (build-only when ctx is None)."""

    def __init__(self, ctx=None):
        self.L = load()
        self.ctx = ctx
        h = C.c_void_p()
        _check(self.L.pz_circuit_create(ctx.h if ctx else None, C.byref(h)))
        self.h = h.value
        self.n_inputs = 0

    def close(self):
        if self.h:
            self.L.pz_circuit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def inputs(self, count):
        w = np.zeros(count, dtype=np.uint32)
        _check(self.L.pz_circuit_inputs(self.h, count, w.ctypes.data))
        self.n_inputs += count
        return w

    def gate(self, op, a, b=0):
        out = C.c_uint32()
        op = GATES.get(op, op) if op != "not" else GATE_NOT
        _check(self.L.pz_circuit_gate(self.h, op, int(a), int(b), C.byref(out)))
        return out.value

    def u8(self, op, a8, b8, sel=None):
        a = np.ascontiguousarray(a8, dtype=np.uint32)
        b = np.ascontiguousarray(b8, dtype=np.uint32)
        out = np.zeros(16, dtype=np.uint32)
        n = C.c_uint32()
        s = C.c_uint32(int(sel)) if sel is not None else None
        _check(self.L.pz_circuit_u8(self.h, U8_OPS[op], a.ctypes.data, b.ctypes.data,
                                    C.byref(s) if s is not None else None, out.ctypes.data,
                                    C.byref(n)))
        return out[:n.value].copy()

    def div_zero_flag(self):
        w = C.c_uint32()
        _check(self.L.pz_circuit_div_zero_flag(self.h, C.byref(w)))
        return w.value

    def stats(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(self.L.pz_circuit_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"wires": a.value, "pbs": b.value, "depth": c.value}

    def level_sizes(self):
        d = self.stats()["depth"]
        out = np.zeros(d + 1, dtype=np.uint32)
        _check(self.L.pz_circuit_level_sizes(self.h, out.ctypes.data, d + 1))
        return out

    def node_levels(self):
        n = self.stats()["wires"]
        out = np.zeros(n, dtype=np.uint32)
        _check(self.L.pz_circuit_node_levels(self.h, out.ctypes.data, n))
        return out

    def nodes(self):
        n = self.stats()["wires"]
        out = np.zeros(4 * n, dtype=np.uint32)
        _check(self.L.pz_circuit_nodes(self.h, out.ctypes.data, n))
        return out.reshape(n, 4)

    def run(self, key, d_inputs):
        _check(self.L.pz_circuit_run(self.h, key.h, d_inputs.ptr if d_inputs is not None else None,
                                     None))

    def read(self, wires):
        w = np.ascontiguousarray(wires, dtype=np.uint32)
        d = self.ctx.array((len(w), self.ctx.N + 1))
        _check(self.L.pz_circuit_read(self.h, w.ctypes.data, len(w), d.ptr, None))
        return d.download()


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("pool_rows", C.c_uint32),
                ("max_batch_pbs", C.c_uint32), ("max_active", C.c_uint32)]


class Server:
    """Continuous-batching request loop over sessions with resident keys."""

    def __init__(self, ctx, hbm_budget_bytes=64 << 30, pool_rows=1 << 16, max_batch_pbs=8192,
                 max_active=4096):
        self.L = load()
        self.ctx = ctx
        cfg = ServerConfig(hbm_budget_bytes, pool_rows, max_batch_pbs, max_active)
        h = C.c_void_p()
        _check(self.L.pz_server_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h.value
        self._keep = []
        self.outs = {}

    def close(self):
        if self.h:
            self.L.pz_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_session(self, rgsw, auto, ksk, ctx=None):
        """A session with this server key; ctx: a context of another parameter
        set (same device and ring size) whose gates run as their own batches."""
        arrs = [np.ascontiguousarray(x, dtype=np.uint64) for x in (rgsw, auto, ksk)]
        self._keep.append(arrs)
        sid = C.c_uint32()
        if ctx is None or ctx is self.ctx:
            _check(self.L.pz_server_add_session(self.h, arrs[0].ctypes.data, arrs[1].ctypes.data,
                                                arrs[2].ctypes.data, C.byref(sid)))
        else:
            self._keep.append(ctx)
            _check(self.L.pz_server_add_session_ctx(self.h, ctx.h, arrs[0].ctypes.data,
                                                    arrs[1].ctypes.data, arrs[2].ctypes.data,
                                                    C.byref(sid)))
        return sid.value

    def submit(self, session, circuit, inputs, out_wires, n_ops, arrival):
        inp = np.ascontiguousarray(inputs, dtype=np.uint64)
        ow = np.ascontiguousarray(out_wires, dtype=np.uint32)
        rid = C.c_uint32()
        _check(self.L.pz_server_submit(self.h, session, circuit.h, inp.ctypes.data, ow.ctypes.data,
                                       len(ow), n_ops, float(arrival), C.byref(rid)))
        self.outs[rid.value] = len(ow)
        return rid.value

    def run(self):
        _check(self.L.pz_server_run(self.h))

    def intake(self, open_):
        _check(self.L.pz_server_intake(self.h, 1 if open_ else 0))

    def result(self, rid):
        out = np.zeros((self.outs[rid], self.ctx.N + 1), dtype=np.uint64)
        _check(self.L.pz_server_result(self.h, rid, out.ctypes.data))
        return out

    def times(self, rid):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        n, k = C.c_uint32(), C.c_uint32()
        _check(self.L.pz_server_request_times(self.h, rid, C.byref(a), C.byref(b), C.byref(c),
                                              C.byref(n), C.byref(k)))
        return {"arrival": a.value, "admitted": b.value, "finished": c.value, "pbs": n.value,
                "ops": k.value}

    def stats(self):
        a, b, c, d = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        w = C.c_double()
        _check(self.L.pz_server_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d),
                                      C.byref(w)))
        return {"total_pbs": a.value, "batches": b.value, "key_uploads": c.value,
                "evictions": d.value, "wall_s": w.value}


class Key:
    """A server key resident in HBM (PbsKey)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle

    def nbytes(self):
        n = C.c_size_t()
        _check(self.ctx.L.pz_key_bytes(self.h, C.byref(n)))
        return n.value

    def free(self):
        if self.h:
            self.ctx.L.pz_key_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class RlweKsk:
    """Ksk(u_j -> s) resident in HBM (ui_to_s_ksk of the server key)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle

    def download(self):
        p = self.ctx.p
        out = np.zeros((2 * p.uitos_d, p.rlwe_n), dtype=np.uint64)
        _check(self.ctx.L.pz_rlwe_ksk_download(self.h, out.ctypes.data))
        return out

    def free(self):
        if self.h:
            self.ctx.L.pz_rlwe_ksk_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One evaluator on one device (BoolEvaluator over the GPU engine)."""

    def __init__(self, params="NI_2P", device=0):
        self.L = load()
        self.p = params_named(params) if isinstance(params, str) else params
        h = C.c_void_p()
        _check(self.L.pz_ctx_create(C.byref(self.p), device, C.byref(h)))
        self.h = h.value
        self.N = self.p.rlwe_n
        self.n = self.p.lwe_n
        self.Q = self.p.rlwe_q

    def close(self):
        if self.h:
            self.L.pz_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        psi, cu, per = C.c_uint64(), C.c_uint32(), C.c_uint32()
        _check(self.L.pz_ctx_info(self.h, C.byref(psi), C.byref(cu), C.byref(per)))
        return {"psi": psi.value, "num_cu": cu.value, "pbs_groups_per_cu": per.value}

    def ntt_tables(self):
        out = np.zeros(4 * self.N + 2, dtype=np.uint64)
        _check(self.L.pz_ctx_ntt_tables(self.h, out.ctypes.data))
        return out

    def array(self, shape, dtype=np.uint64):
        return DeviceArray(self, shape, dtype)

    def to_device(self, host, dtype=None):
        a = np.ascontiguousarray(host, dtype=dtype or host.dtype)
        return DeviceArray(self, a.shape, a.dtype).upload(a)

    def sync(self):
        _check(self.L.pz_sync(self.h))

    # ---- operators ----
    def ntt(self, which, d_rows, nrows):
        f = {"forward": self.L.pz_ntt_forward, "forward_lazy": self.L.pz_ntt_forward_lazy,
             "backward": self.L.pz_ntt_backward,
             "backward_lazy": self.L.pz_ntt_backward_lazy}[which]
        _check(f(self.h, d_rows.ptr, nrows, None))

    def shoup_matrix_fma(self, d_out, d_a, d_as, d_b, d, n):
        _check(self.L.pz_shoup_matrix_fma(self.h, d_out.ptr, d_a.ptr, d_as.ptr, d_b.ptr, d, n,
                                          None))

    def decompose(self, which, d_in, d_digits, count):
        _check(self.L.pz_decompose(self.h, which, d_in.ptr, d_digits.ptr, count, None))

    def key_upload(self, rgsw, auto, lwe_ksk):
        rgsw = np.ascontiguousarray(rgsw, dtype=np.uint64)
        auto = np.ascontiguousarray(auto, dtype=np.uint64)
        ksk = np.ascontiguousarray(lwe_ksk, dtype=np.uint64)
        h = C.c_void_p()
        _check(self.L.pz_key_upload(self.h, rgsw.ctypes.data, auto.ctypes.data,
                                    ksk.ctypes.data, C.byref(h)))
        return Key(self, h.value)

    # ---- seeded client-side keys and encryption ----
    @staticmethod
    def seed32(seed):
        """32-byte seed array.  Bind it to a name before passing its address
        to C: a temporary's buffer is freed before the call reads it."""
        if isinstance(seed, (bytes, bytearray)):
            return np.frombuffer(bytes(seed).ljust(32, b"\0")[:32], dtype=np.uint8).copy()
        return np.frombuffer(int(seed).to_bytes(32, "little"), dtype=np.uint8).copy()

    def keygen_secrets(self, seed, parties):
        s = np.zeros(self.N, dtype=np.int32)
        z = np.zeros(self.n, dtype=np.int32)
        sd_b = self.seed32(seed)
        _check(self.L.pz_keygen_secrets(self.h, sd_b.ctypes.data, parties,
                                        s.ctypes.data, z.ctypes.data))
        return s, z

    def keygen_server(self, seed, s, z):
        s = np.ascontiguousarray(s, dtype=np.int32)
        z = np.ascontiguousarray(z, dtype=np.int32)
        h = C.c_void_p()
        sd_b = self.seed32(seed)
        _check(self.L.pz_keygen_server(self.h, sd_b.ctypes.data, s.ctypes.data,
                                       z.ctypes.data, C.byref(h)))
        return Key(self, h.value)

    def key_download(self, key):
        p = self.p
        rgsw = np.zeros((p.lwe_n, 2 * p.rlrg_da + 2 * p.rlrg_db, p.rlwe_n), dtype=np.uint64)
        auto = np.zeros((p.w + 1, 2 * p.auto_d, p.rlwe_n), dtype=np.uint64)
        ksk = np.zeros((p.rlwe_n * p.lwe_d, p.lwe_n + 1), dtype=np.uint64)
        _check(self.L.pz_key_download(key.h, rgsw.ctypes.data, auto.ctypes.data,
                                      ksk.ctypes.data))
        return rgsw, auto, ksk

    def encrypt_bools(self, s, bits, seed):
        s = np.ascontiguousarray(s, dtype=np.int32)
        b = np.ascontiguousarray(bits, dtype=np.uint8)
        out = np.zeros((len(b), len(s) + 1), dtype=np.uint64)
        sd_b = self.seed32(seed)
        _check(self.L.pz_encrypt_bools(self.h, s.ctypes.data, len(s), sd_b.ctypes.data,
                                       b.ctypes.data, len(b), out.ctypes.data))
        return out

    def decrypt_bools(self, s, cts):
        s = np.ascontiguousarray(s, dtype=np.int32)
        c = np.ascontiguousarray(cts, dtype=np.uint64)
        out = np.zeros(c.shape[0], dtype=np.uint8)
        _check(self.L.pz_decrypt_bools(self.h, s.ctypes.data, len(s), c.ctypes.data, c.shape[0],
                                       out.ctypes.data))
        return out

    # ---- input path (non-interactive sets) ----
    def keygen_user_secret(self, seed):
        u = np.zeros(self.N, dtype=np.int32)
        sd_b = self.seed32(seed)
        _check(self.L.pz_keygen_user_secret(self.h, sd_b.ctypes.data, u.ctypes.data))
        return u

    def keygen_rlwe_ksk(self, seed, u, s):
        u = np.ascontiguousarray(u, dtype=np.int32)
        s = np.ascontiguousarray(s, dtype=np.int32)
        h = C.c_void_p()
        sd_b = self.seed32(seed)
        _check(self.L.pz_keygen_rlwe_ksk(self.h, sd_b.ctypes.data, u.ctypes.data,
                                         s.ctypes.data, C.byref(h)))
        return RlweKsk(self, h.value)

    def rlwe_ksk_upload(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        h = C.c_void_p()
        _check(self.L.pz_rlwe_ksk_upload(self.h, rows.ctypes.data, C.byref(h)))
        return RlweKsk(self, h.value)

    def encrypt_rlwe_bools(self, u, bits, seed):
        u = np.ascontiguousarray(u, dtype=np.int32)
        b = np.ascontiguousarray(bits, dtype=np.uint8)
        nr = (len(b) + self.N - 1) // self.N
        out = np.zeros((nr, 2, self.N), dtype=np.uint64)
        sd_b = self.seed32(seed)
        _check(self.L.pz_encrypt_rlwe_bools(self.h, u.ctypes.data, sd_b.ctypes.data,
                                            b.ctypes.data, len(b), out.ctypes.data))
        return out

    def rlwe_key_switch(self, ksk, d_in, d_out, B):
        _check(self.L.pz_rlwe_key_switch(self.h, ksk.h, d_in.ptr, d_out.ptr, B, None))

    def sample_extract(self, d_rlwe, d_src, count, d_lwe):
        _check(self.L.pz_sample_extract(self.h, d_rlwe.ptr, d_src.ptr, count, d_lwe.ptr, None))

    # ---- non-interactive server key: shares (client) and aggregation (server)
    @property
    def ni_dmax(self):
        return max(self.p.rgrg_da, self.p.rgrg_db)

    def ni_share_uitos(self, crs, noise_seed, me, total, s, u):
        d, N = self.p.uitos_d, self.N
        s = np.ascontiguousarray(s, dtype=np.int32)
        u = np.ascontiguousarray(u, dtype=np.int32)
        ksk_b = np.zeros((d, N), dtype=np.uint64)
        zo = np.zeros((max(total - 1, 1), d, N), dtype=np.uint64)
        crs_b = self.seed32(crs)
        ns_b = self.seed32(noise_seed)
        _check(self.L.pz_ni_share_uitos(self.h, crs_b.ctypes.data,
                                        ns_b.ctypes.data, me, total,
                                        s.ctypes.data, u.ctypes.data, ksk_b.ctypes.data,
                                        zo.ctypes.data))
        return ksk_b, zo[:total - 1]

    def ni_share_rgsw(self, crs, noise_seed, me, total, s, u, z, l0, count):
        N = self.N
        s = np.ascontiguousarray(s, dtype=np.int32)
        u = np.ascontiguousarray(u, dtype=np.int32)
        z = np.ascontiguousarray(z, dtype=np.int32)
        zero = np.zeros((count, self.p.rgrg_da, N), dtype=np.uint64)
        cts = np.zeros((count, self.ni_dmax, N), dtype=np.uint64)
        crs_b = self.seed32(crs)
        ns_b = self.seed32(noise_seed)
        _check(self.L.pz_ni_share_rgsw(self.h, crs_b.ctypes.data,
                                       ns_b.ctypes.data, me, total,
                                       s.ctypes.data, u.ctypes.data, z.ctypes.data, l0, count,
                                       zero.ctypes.data, cts.ctypes.data))
        return zero, cts

    def ni_uitos_ksk(self, crs, total, user, d_ksk_b, d_zero_for_user):
        h = C.c_void_p()
        crs_b = self.seed32(crs)
        _check(self.L.pz_ni_uitos_ksk(self.h, crs_b.ctypes.data, total, user,
                                      d_ksk_b.ptr,
                                      d_zero_for_user.ptr if d_zero_for_user is not None else None,
                                      C.byref(h)))
        return RlweKsk(self, h.value)

    def ni_aggregate_rgsw(self, crs, total, ksks, l0, count, d_zero, d_cts, d_out):
        arr = (C.c_void_p * len(ksks))(*[k.h for k in ksks])
        crs_b = self.seed32(crs)
        _check(self.L.pz_ni_aggregate_rgsw(self.h, crs_b.ctypes.data, total, arr, l0,
                                           count, d_zero.ptr, d_cts.ptr, d_out.ptr, None))

    def ni_share_auto_ksk(self, crs, noise_seed, s, z):
        """A party's automorphism-key b parts [(w+1)][d_auto][N] and LWE-KSK
        b parts [N * d_ks]."""
        p, N = self.p, self.N
        s = np.ascontiguousarray(s, dtype=np.int32)
        z = np.ascontiguousarray(z, dtype=np.int32)
        auto_b = np.zeros((p.w + 1, p.auto_d, N), dtype=np.uint64)
        ksk_b = np.zeros(N * p.lwe_d, dtype=np.uint64)
        crs_b, ns_b = self.seed32(crs), self.seed32(noise_seed)
        _check(self.L.pz_ni_share_auto_ksk(self.h, crs_b.ctypes.data, ns_b.ctypes.data,
                                           s.ctypes.data, z.ctypes.data, auto_b.ctypes.data,
                                           ksk_b.ctypes.data))
        return auto_b, ksk_b

    def ni_aggregate_auto_ksk(self, total, d_auto_shares, d_ksk_shares, d_auto_b, d_ksk_b):
        _check(self.L.pz_ni_aggregate_auto_ksk(self.h, total, d_auto_shares.ptr,
                                               d_ksk_shares.ptr, d_auto_b.ptr, d_ksk_b.ptr, None))

    def ni_expand_server_key(self, crs, d_rgsw, d_auto_b, d_ksk_b):
        h = C.c_void_p()
        crs_b = self.seed32(crs)
        _check(self.L.pz_ni_expand_server_key(self.h, crs_b.ctypes.data, d_rgsw.ptr,
                                              d_auto_b.ptr, d_ksk_b.ptr, C.byref(h)))
        return Key(self, h.value)

    def ni_key_share(self, crs, noise_seed, me, total, s, u, z):
        """Party `me`'s whole share as a wire-format blob (uint8 array)."""
        s = np.ascontiguousarray(s, dtype=np.int32)
        u = np.ascontiguousarray(u, dtype=np.int32)
        z = np.ascontiguousarray(z, dtype=np.int32)
        out = aligned_bytes(ni_share_bytes(self.p, me, total))
        crs_b, ns_b = self.seed32(crs), self.seed32(noise_seed)
        _check(self.L.pz_ni_key_share(self.h, crs_b.ctypes.data, ns_b.ctypes.data, me, total,
                                      s.ctypes.data, u.ctypes.data, z.ctypes.data,
                                      out.ctypes.data))
        return out

    def ni_aggregate_shares(self, shares):
        """The seeded server key blob of a list of share blobs."""
        total = len(shares)
        keep = []
        for b in shares:
            if b.ctypes.data % 8:
                a = aligned_bytes(len(b))
                a[:] = b
                b = a
            keep.append(b)
        arr = (C.c_void_p * total)(*[b.ctypes.data for b in keep])
        out = aligned_bytes(ni_seeded_key_bytes(self.p, total))
        _check(self.L.pz_ni_aggregate_shares(self.h, arr, total, out.ctypes.data))
        return out

    def ni_server_key_from_seeded(self, blob, total=None):
        """(Key, [RlweKsk] or None) from a seeded server key blob; with
        total given, the parties' Ksk(u_j -> s) are made too."""
        h = C.c_void_p()
        ks = (C.c_void_p * total)() if total else None
        _check(self.L.pz_ni_server_key_from_seeded(self.h, blob.ctypes.data, C.byref(h), ks))
        return Key(self, h.value), ([RlweKsk(self, k) for k in ks] if total else None)

    def autotune(self, key, B):
        """Pick the fastest blind-rotation launch shape for batches like B."""
        out = np.zeros(3, dtype=np.uint32)
        _check(self.L.pz_autotune(self.h, key.h, B, out.ctypes.data))
        names = {0: "one_half", 1: "split_latency", 2: "split", 3: "split4_latency",
                 4: "split4"}
        return {"variant": names[int(out[0])], "groups_per_cu": int(out[1]),
                "ms_per_batch": int(out[2]) / 1e3}

    def rgsw_by_rgsw(self, d_rgsw0, d_rgsw1, B):
        _check(self.L.pz_rgsw_by_rgsw(self.h, d_rgsw0.ptr, d_rgsw1.ptr, B, None))

    def rlwe_by_rgsw(self, key, d_rlwe, d_idx, d_triv, B):
        _check(self.L.pz_rlwe_by_rgsw(self.h, key.h, d_rlwe.ptr, d_idx.ptr, d_triv.ptr, B, None))

    def rlwe_auto(self, key, d_rlwe, d_idx, d_triv, B):
        _check(self.L.pz_rlwe_auto(self.h, key.h, d_rlwe.ptr, d_idx.ptr, d_triv.ptr, B, None))

    def lwe_key_switch(self, key, d_in, d_out, B):
        _check(self.L.pz_lwe_key_switch(self.h, key.h, d_in.ptr, d_out.ptr, B, None))

    def pbs(self, key, d_gate, d_lwe, B):
        _check(self.L.pz_pbs(self.h, key.h, d_gate.ptr, d_lwe.ptr, B, None))

    def gates(self, key, d_ops, d_c0, d_c1, d_out, B):
        _check(self.L.pz_gates(self.h, key.h, d_ops.ptr, d_c0.ptr, d_c1.ptr, d_out.ptr, B, None))

    def not_(self, d_c, B):
        _check(self.L.pz_not(self.h, d_c.ptr, B, None))

    def last_auto_counts(self, B):
        out = np.zeros(B, dtype=np.uint32)
        _check(self.L.pz_last_auto_counts(self.h, out.ctypes.data, B))
        return out

    def set_timing(self, on):
        _check(self.L.pz_set_timing(self.h, 1 if on else 0))

    def set_blind_rotate_mode(self, mode):
        """0 automatic, 1 one work group per bootstrap, 2 two half groups,
        3 two half groups with 4 coefficients per thread."""
        _check(self.L.pz_set_blind_rotate_mode(self.h, mode))

    def valu_probe(self, kind, iters=4096):
        """Measured rate of one VALU instruction form (kind 0-3, lane ops/s)
        or of the Shoup multiply (kind 4, multiplies/s)."""
        v = C.c_double()
        _check(self.L.pz_valu_probe(self.h, kind, iters, C.byref(v)))
        return v.value

    def last_timing(self):
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        _check(self.L.pz_last_timing(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"prep_ms": a.value, "ks_ms": b.value, "blind_rotate_ms": c.value}
