"""ctypes front end of the C oracle (ORACLE / TEST INFRASTRUCTURE ONLY).

The oracle restates phantom-zone's PBS path on the CPU; see pz_oracle.c for
the reference citations and for how each function is pinned.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

PARAM_SETS = {
    "NI_2P": 0, "NI_4P_HB_FR": 1, "NI_8P": 2, "SP_TEST": 3,
    "I_2P_LB_SR": 4, "I_4P": 5, "I_8P": 6, "NI_4P_LB_SR": 7,
}
GATES = {"nand": 0, "and": 1, "or": 2, "nor": 3, "xor": 4, "xnor": 5}

u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64).reshape(-1)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32).reshape(-1)


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

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class KeySizes(C.Structure):
    _fields_ = [(n, C.c_size_t) for n in
                ("rgsw_rows", "rgsw_words", "auto_rows", "auto_words", "ksk_rows", "ksk_words")]


class Key(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("rgsw", "rgsw_shoup", "auto_key", "auto_shoup", "lwe_ksk")]


class PbsStats(C.Structure):
    _fields_ = [("ext_products", C.c_uint32), ("autos", C.c_uint32),
                ("trivial_autos", C.c_uint32)]


def build():
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "pzo_params_named": (C.c_int, [C.c_int, C.POINTER(Params)]),
            "pzo_chacha_block": (None, [u8p, C.c_uint64, C.c_uint64, C.c_int, u32p]),
            "pzo_rng_new": (vp, [u8p]),
            "pzo_rng_free": (None, [vp]),
            "pzo_rng_next_u32": (C.c_uint32, [vp]),
            "pzo_rng_next_u64": (C.c_uint64, [vp]),
            "pzo_rng_uniform": (C.c_uint64, [vp, C.c_uint64]),
            "pzo_rng_gauss_round": (C.c_int64, [vp]),
            "pzo_rng_gauss_trunc": (C.c_int64, [vp]),
            "pzo_ternary_hw": (None, [vp, i32p, C.c_uint32, C.c_uint32]),
            "pzo_shoup_repr": (C.c_uint64, [C.c_uint64, C.c_uint64]),
            "pzo_shoup_mul": (C.c_uint64, [C.c_uint64] * 4),
            "pzo_mul_mod_fast": (C.c_uint64, [C.c_uint64] * 3),
            "pzo_mul_mod_fast_lazy": (C.c_uint64, [C.c_uint64] * 3),
            "pzo_mod_exponent": (C.c_uint64, [C.c_uint64] * 3),
            "pzo_shoup_matrix_fma": (None, [C.c_uint64, u64p, u64p, u64p, u64p,
                                            C.c_uint32, C.c_uint32]),
            "pzo_ntt_new": (vp, [C.c_uint64, C.c_uint32]),
            "pzo_ntt_free": (None, [vp]),
            "pzo_ntt_psi": (C.c_uint64, [vp]),
            "pzo_ntt_tables": (None, [vp, u64p]),
            "pzo_ntt_rows": (None, [vp, C.c_int, u64p, C.c_uint32]),
            "pzo_negacyclic_mul_schoolbook": (None, [C.c_uint64, u64p, u64p, u64p, C.c_uint32]),
            "pzo_decompose_iter": (None, [C.c_uint64, C.c_uint32, C.c_uint32, u64p, u64p,
                                          C.c_size_t]),
            "pzo_decompose_to_vec": (None, [C.c_uint64, C.c_uint32, C.c_uint32, u64p, u64p,
                                            C.c_size_t]),
            "pzo_round_value": (C.c_uint64, [C.c_uint64, C.c_uint32]),
            "pzo_recompose": (C.c_uint64, [C.c_uint64, C.c_uint32, C.c_uint32, u64p]),
            "pzo_decomposer_ignore_bits": (C.c_uint32, [C.c_uint64, C.c_uint32, C.c_uint32]),
            "pzo_eval_new": (vp, [C.POINTER(Params)]),
            "pzo_eval_free": (None, [vp]),
            "pzo_eval_ntt": (vp, [vp]),
            "pzo_eval_dlog_map": (None, [vp, u32p]),
            "pzo_eval_auto_element": (C.c_int64, [vp, C.c_uint32]),
            "pzo_eval_auto_map": (None, [vp, C.c_uint32, u32p, u8p]),
            "pzo_eval_test_vec": (None, [vp, C.c_uint32, u64p]),
            "pzo_generate_auto_map": (None, [C.c_uint32, C.c_int64, u32p, u8p]),
            "pzo_key_sizes_get": (None, [vp, C.POINTER(KeySizes)]),
            "pzo_keygen_secrets": (None, [vp, u8p, C.c_uint32, i32p, i32p]),
            "pzo_keygen_server": (None, [vp, u8p, i32p, i32p, C.POINTER(Key)]),
            "pzo_key_shoup": (None, [vp, C.POINTER(Key)]),
            "pzo_rlwe_by_rgsw_shoup": (None, [vp, u64p, u64p, u64p, C.c_int]),
            "pzo_rlwe_by_rgsw": (None, [vp, u64p, u64p, C.c_int]),
            "pzo_rlwe_auto_shoup": (None, [vp, u64p, u64p, u64p, C.c_uint32, C.c_int]),
            "pzo_rlwe_auto": (None, [vp, u64p, u64p, C.c_uint32, C.c_int]),
            "pzo_lwe_key_switch": (None, [vp, u64p, u64p, u64p]),
            "pzo_moddown": (C.c_uint64, [vp, C.c_uint64]),
            "pzo_mod_switch_odd": (C.c_uint32, [C.c_double, C.c_double, C.c_double]),
            "pzo_sample_extract": (None, [vp, u64p, u64p, C.c_uint32]),
            "pzo_monomial_mul": (None, [C.c_uint64, u64p, u64p, C.c_uint32, C.c_int,
                                        C.c_uint32]),
            "pzo_pbs": (None, [vp, C.POINTER(Key), C.c_uint32, u64p, C.POINTER(PbsStats)]),
            "pzo_br_schedule": (C.c_uint32, [vp, u64p, u32p, C.c_uint32]),
            "pzo_gate": (None, [vp, C.POINTER(Key), C.c_uint32, u64p, u64p]),
            "pzo_not": (None, [vp, u64p]),
            "pzo_pbs_batch_threads": (C.c_double, [vp, C.POINTER(Key), C.c_uint32, u64p,
                                                   C.c_uint32, C.c_uint32]),
            "pzo_gates_batch_threads": (C.c_double, [vp, C.POINTER(Key), u32p, u64p, u64p,
                                                     C.c_uint32, C.c_uint32]),
            "pzo_encrypt_lwe": (None, [vp, u8p, i32p, C.c_uint32, C.c_uint64, u64p]),
            "pzo_keygen_user_secret": (None, [vp, u8p, i32p]),
            "pzo_keygen_rlwe_ksk": (None, [vp, u8p, i32p, i32p, u64p]),
            "pzo_rlwe_key_switch": (None, [vp, u64p, u64p, u64p, u64p]),
            "pzo_encrypt_rlwe_bools": (None, [vp, u8p, i32p, u8p, C.c_size_t, u64p]),
            "pzo_rgsw_by_rgsw": (None, [vp, u64p, u64p]),
            "pzo_puncture": (None, [u8p, C.c_uint32, u8p]),
            "pzo_ni_leader": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
            "pzo_ni_share_uitos": (None, [vp, u8p, u8p, C.c_uint32, C.c_uint32, i32p, i32p,
                                          u64p, u64p]),
            "pzo_ni_share_rgsw": (None, [vp, u8p, u8p, C.c_uint32, C.c_uint32, i32p, i32p, i32p,
                                         C.c_uint32, C.c_uint32, u64p, u64p]),
            "pzo_ni_uitos_ksk": (None, [vp, u8p, C.c_uint32, C.c_uint32, u64p, u64p, u64p]),
            "pzo_ni_produce": (None, [vp, u8p, C.c_uint32, C.c_uint32, u64p, u64p, u64p,
                                      C.c_int, u64p]),
            "pzo_ni_aggregate_rgsw": (None, [vp, u8p, C.c_uint32, u64p, C.c_uint32, C.c_uint32,
                                             u64p, u64p, u64p]),
            "pzo_ni_share_auto_ksk": (None, [vp, u8p, u8p, i32p, i32p, u64p, u64p]),
            "pzo_ni_aggregate_auto_ksk": (None, [vp, C.c_uint32, u64p, u64p, u64p, u64p]),
            "pzo_ni_expand_server_key": (None, [vp, u8p, u64p, u64p, u64p, C.POINTER(Key)]),
            "pzo_decrypt_lwe": (C.c_uint64, [vp, u64p, i32p, C.c_uint32]),
            "pzo_decode_bool": (C.c_int, [vp, C.c_uint64]),
            "pzo_true_el": (C.c_uint64, [vp]),
            "pzo_false_el": (C.c_uint64, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def seed_bytes(seed):
    """32-byte seed array from an int or bytes."""
    if isinstance(seed, np.ndarray):
        seed = seed.astype(np.uint8).tobytes()
    if isinstance(seed, (bytes, bytearray)):
        b = bytes(seed).ljust(32, b"\0")[:32]
    else:
        b = int(seed).to_bytes(32, "little")
    return np.frombuffer(b, dtype=np.uint8).copy()


class OracleKey:
    """Server key in the reference's layout (numpy arrays owned here)."""

    def __init__(self, sizes, N):
        self.rgsw = np.zeros(sizes.rgsw_words, dtype=np.uint64)
        self.rgsw_shoup = np.zeros(sizes.rgsw_words, dtype=np.uint64)
        self.auto_key = np.zeros(sizes.auto_words, dtype=np.uint64)
        self.auto_shoup = np.zeros(sizes.auto_words, dtype=np.uint64)
        self.lwe_ksk = np.zeros(sizes.ksk_words, dtype=np.uint64)
        self.sizes = sizes
        self.N = N

    def cstruct(self):
        return Key(self.rgsw.ctypes.data, self.rgsw_shoup.ctypes.data,
                   self.auto_key.ctypes.data, self.auto_shoup.ctypes.data,
                   self.lwe_ksk.ctypes.data)

    def rgsw_ct(self, i):
        rows = self.sizes.rgsw_rows
        n = rows * self.N
        return (self.rgsw[i * n:(i + 1) * n].reshape(rows, self.N),
                self.rgsw_shoup[i * n:(i + 1) * n].reshape(rows, self.N))

    def auto_ct(self, k):
        rows = self.sizes.auto_rows
        n = rows * self.N
        return (self.auto_key[k * n:(k + 1) * n].reshape(rows, self.N),
                self.auto_shoup[k * n:(k + 1) * n].reshape(rows, self.N))


class Oracle:
    """One evaluator (BoolEvaluator::new restated) for a parameter set."""

    def __init__(self, params="NI_2P"):
        L = lib()
        self.L = L
        p = Params()
        if isinstance(params, str):
            assert L.pzo_params_named(PARAM_SETS[params], C.byref(p)) == 0
        else:
            p = params
        self.p = p
        self.name = params if isinstance(params, str) else "custom"
        self.ev = L.pzo_eval_new(C.byref(p))
        self.ntt = L.pzo_eval_ntt(self.ev)
        self.N = p.rlwe_n
        self.n = p.lwe_n
        self.Q = p.rlwe_q
        ks = KeySizes()
        L.pzo_key_sizes_get(self.ev, C.byref(ks))
        self.sizes = ks

    def __del__(self):
        try:
            self.L.pzo_eval_free(self.ev)
        except Exception:
            pass

    # ---- NTT ----
    def ntt_rows(self, which, rows):
        """which: 'forward_lazy', 'forward', 'backward_lazy', 'backward'."""
        code = {"forward_lazy": 0, "forward": 1, "backward_lazy": 2, "backward": 3}[which]
        a = np.ascontiguousarray(rows, dtype=np.uint64).copy()
        nrows = a.size // self.N
        self.L.pzo_ntt_rows(self.ntt, code, a.reshape(-1), nrows)
        return a

    def ntt_tables(self):
        out = np.zeros(4 * self.N + 2, dtype=np.uint64)
        self.L.pzo_ntt_tables(self.ntt, out)
        return out

    def psi(self):
        return int(self.L.pzo_ntt_psi(self.ntt))

    # ---- evaluator constants ----
    def dlog_map(self):
        out = np.zeros(self.p.br_q, dtype=np.uint32)
        self.L.pzo_eval_dlog_map(self.ev, out)
        return out

    def auto_map(self, k):
        idx = np.zeros(self.N, dtype=np.uint32)
        sg = np.zeros(self.N, dtype=np.uint8)
        self.L.pzo_eval_auto_map(self.ev, k, idx, sg)
        return idx, sg

    def test_vec(self, gate):
        out = np.zeros(self.p.br_q // 2, dtype=np.uint64)
        self.L.pzo_eval_test_vec(self.ev, GATES.get(gate, gate), out)
        return out

    # ---- keys ----
    def keygen(self, seed_secret=0, seed_key=1, parties=2):
        s = np.zeros(self.N, dtype=np.int32)
        z = np.zeros(self.n, dtype=np.int32)
        self.L.pzo_keygen_secrets(self.ev, seed_bytes(seed_secret), parties, s, z)
        key = OracleKey(self.sizes, self.N)
        self.L.pzo_keygen_server(self.ev, seed_bytes(seed_key), s, z, C.byref(key.cstruct()))
        return s, z, key

    # ---- runtime ops ----
    def rlwe_by_rgsw_shoup(self, rlwe, rgsw, rgsw_shoup, trivial=False):
        r = np.ascontiguousarray(rlwe, dtype=np.uint64).reshape(-1).copy()
        self.L.pzo_rlwe_by_rgsw_shoup(self.ev, r, np.ascontiguousarray(rgsw).reshape(-1),
                                      np.ascontiguousarray(rgsw_shoup).reshape(-1), int(trivial))
        return r.reshape(2, self.N)

    def rlwe_by_rgsw(self, rlwe, rgsw, trivial=False):
        r = np.ascontiguousarray(rlwe, dtype=np.uint64).reshape(-1).copy()
        self.L.pzo_rlwe_by_rgsw(self.ev, r, np.ascontiguousarray(rgsw).reshape(-1), int(trivial))
        return r.reshape(2, self.N)

    def rlwe_auto_shoup(self, rlwe, ksk, ksk_shoup, k, trivial=False):
        r = np.ascontiguousarray(rlwe, dtype=np.uint64).reshape(-1).copy()
        self.L.pzo_rlwe_auto_shoup(self.ev, r, np.ascontiguousarray(ksk).reshape(-1),
                                   np.ascontiguousarray(ksk_shoup).reshape(-1), k, int(trivial))
        return r.reshape(2, self.N)

    def rlwe_auto(self, rlwe, ksk, k, trivial=False):
        r = np.ascontiguousarray(rlwe, dtype=np.uint64).reshape(-1).copy()
        self.L.pzo_rlwe_auto(self.ev, r, np.ascontiguousarray(ksk).reshape(-1), k, int(trivial))
        return r.reshape(2, self.N)

    def lwe_key_switch(self, lwe_in, ksk):
        out = np.zeros(self.n + 1, dtype=np.uint64)
        self.L.pzo_lwe_key_switch(self.ev, out, np.ascontiguousarray(lwe_in, dtype=np.uint64),
                                  np.ascontiguousarray(ksk, dtype=np.uint64))
        return out

    def moddown(self, v):
        return int(self.L.pzo_moddown(self.ev, int(v)))

    def pbs(self, key, gate, lwe):
        out = np.ascontiguousarray(lwe, dtype=np.uint64).copy()
        st = PbsStats()
        self.L.pzo_pbs(self.ev, C.byref(key.cstruct()), GATES.get(gate, gate), out, C.byref(st))
        return out, st

    def br_schedule(self, lwe_ks, cap=8192):
        ops = np.zeros(cap, dtype=np.uint32)
        n = self.L.pzo_br_schedule(self.ev, np.ascontiguousarray(lwe_ks, dtype=np.uint64), ops, cap)
        return ops[:n]

    def gate(self, key, gate, c0, c1):
        out = np.ascontiguousarray(c0, dtype=np.uint64).copy()
        self.L.pzo_gate(self.ev, C.byref(key.cstruct()), GATES.get(gate, gate), out,
                        np.ascontiguousarray(c1, dtype=np.uint64))
        return out

    def not_(self, c0):
        out = np.ascontiguousarray(c0, dtype=np.uint64).copy()
        self.L.pzo_not(self.ev, out)
        return out

    def pbs_batch_threads(self, key, gate, lwes, threads):
        arr = np.ascontiguousarray(lwes, dtype=np.uint64)
        count = arr.shape[0]
        secs = self.L.pzo_pbs_batch_threads(self.ev, C.byref(key.cstruct()),
                                            GATES.get(gate, gate), arr.reshape(-1), count, threads)
        return arr, secs

    def gates_batch(self, key, gates, c0, c1, threads=None):
        """gates[i] of (c0[i], c1[i]) for every i, on `threads` host threads
        (default: every core); the same words as gate() on each item."""
        out = np.ascontiguousarray(c0, dtype=np.uint64).copy()
        b = np.ascontiguousarray(c1, dtype=np.uint64)
        g = np.ascontiguousarray(gates, dtype=np.uint32)
        count = out.shape[0]
        if count == 0:
            return out
        threads = threads or os.cpu_count() or 1
        self.L.pzo_gates_batch_threads(self.ev, C.byref(key.cstruct()), g, out.reshape(-1),
                                       b.reshape(-1), count, min(threads, count))
        return out

    # ---- input path (non-interactive sets) ----
    def keygen_user_secret(self, seed):
        u = np.zeros(self.N, dtype=np.int32)
        self.L.pzo_keygen_user_secret(self.ev, seed_bytes(seed), u)
        return u

    def keygen_rlwe_ksk(self, seed, u, s):
        ksk = np.zeros((2 * self.p.uitos_d, self.N), dtype=np.uint64)
        self.L.pzo_keygen_rlwe_ksk(self.ev, seed_bytes(seed), np.ascontiguousarray(u, dtype=np.int32),
                                   np.ascontiguousarray(s, dtype=np.int32), ksk.reshape(-1))
        return ksk

    def shoup_rows(self, rows):
        """Shoup twins of a matrix of residues (ToShoup, src/utils.rs:49-73)."""
        q = self.Q
        return np.array([[(int(v) << 64) // q for v in row] for row in np.asarray(rows)],
                        dtype=np.uint64)

    def rlwe_key_switch(self, rlwe, ksk, ksk_shoup):
        out = np.zeros((2, self.N), dtype=np.uint64)
        self.L.pzo_rlwe_key_switch(self.ev, np.ascontiguousarray(rlwe, dtype=np.uint64).reshape(-1),
                                   np.ascontiguousarray(ksk, dtype=np.uint64).reshape(-1),
                                   np.ascontiguousarray(ksk_shoup, dtype=np.uint64).reshape(-1),
                                   out.reshape(-1))
        return out

    def encrypt_rlwe_bools(self, u, bits, seed):
        b = np.ascontiguousarray(bits, dtype=np.uint8)
        nr = (len(b) + self.N - 1) // self.N
        out = np.zeros((nr, 2, self.N), dtype=np.uint64)
        self.L.pzo_encrypt_rlwe_bools(self.ev, seed_bytes(seed), np.ascontiguousarray(u, dtype=np.int32),
                                      b, len(b), out.reshape(-1))
        return out

    def sample_extract(self, rlwe, index):
        out = np.zeros(self.N + 1, dtype=np.uint64)
        self.L.pzo_sample_extract(self.ev, out, np.ascontiguousarray(rlwe, dtype=np.uint64).reshape(-1),
                                  index)
        return out

    def rgsw_by_rgsw(self, rgsw0, rgsw1):
        out = np.ascontiguousarray(rgsw0, dtype=np.uint64).copy()
        self.L.pzo_rgsw_by_rgsw(self.ev, out.reshape(-1),
                                np.ascontiguousarray(rgsw1, dtype=np.uint64).reshape(-1))
        return out

    # ---- non-interactive server key: shares and aggregation ----
    @property
    def embedding(self):
        return 2 * self.N // self.p.br_q

    @property
    def ni_dmax(self):
        return max(self.p.rgrg_da, self.p.rgrg_db)

    def puncture(self, seed, times):
        out = np.zeros(32, dtype=np.uint8)
        self.L.pzo_puncture(seed_bytes(seed), times, out)
        return out

    def keygen_party(self, seed):
        """One party's (s_j, z_j): the secrets of a single-party draw."""
        s = np.zeros(self.N, dtype=np.int32)
        z = np.zeros(self.p.lwe_n, dtype=np.int32)
        self.L.pzo_keygen_secrets(self.ev, seed_bytes(seed), 1, s, z)
        return s, z

    def ni_leader(self, l, total):
        return int(self.L.pzo_ni_leader(l, total, self.p.lwe_n))

    def ni_share_uitos(self, crs, noise_seed, me, total, s, u):
        d, N = self.p.uitos_d, self.N
        ksk_b = np.zeros((d, N), dtype=np.uint64)
        zo = np.zeros((max(total - 1, 1), d, N), dtype=np.uint64)
        self.L.pzo_ni_share_uitos(self.ev, seed_bytes(crs), seed_bytes(noise_seed), me, total,
                                  _i32(s), _i32(u), ksk_b.reshape(-1), zo.reshape(-1))
        return ksk_b, zo[:total - 1]

    def ni_share_rgsw(self, crs, noise_seed, me, total, s, u, z, l0, count):
        N = self.N
        zero = np.zeros((count, self.p.rgrg_da, N), dtype=np.uint64)
        cts = np.zeros((count, self.ni_dmax, N), dtype=np.uint64)
        self.L.pzo_ni_share_rgsw(self.ev, seed_bytes(crs), seed_bytes(noise_seed), me, total,
                                 _i32(s), _i32(u), _i32(z), l0, count, zero.reshape(-1),
                                 cts.reshape(-1))
        return zero, cts

    def ni_uitos_ksk(self, crs, total, user, ksk_b, zero_for_user):
        d, N = self.p.uitos_d, self.N
        out = np.zeros((2 * d, N), dtype=np.uint64)
        zf = np.ascontiguousarray(zero_for_user, dtype=np.uint64).reshape(-1)
        if zf.size == 0:
            zf = np.zeros(1, dtype=np.uint64)
        self.L.pzo_ni_uitos_ksk(self.ev, seed_bytes(crs), total, user, _u64(ksk_b), zf,
                                out.reshape(-1))
        return out

    def ni_produce(self, crs, total, l, ksk, zero_l, ct, leader):
        p = self.p
        rows = 2 * (p.rlrg_da + p.rlrg_db) if leader else 2 * (p.rgrg_da + p.rgrg_db)
        out = np.zeros((rows, self.N), dtype=np.uint64)
        self.L.pzo_ni_produce(self.ev, seed_bytes(crs), total, l, _u64(ksk), _u64(zero_l),
                              _u64(ct), 1 if leader else 0, out.reshape(-1))
        return out

    def ni_aggregate_rgsw(self, crs, total, ksks, l0, count, zero, cts):
        p = self.p
        out = np.zeros((count, 2 * (p.rlrg_da + p.rlrg_db), self.N), dtype=np.uint64)
        self.L.pzo_ni_aggregate_rgsw(self.ev, seed_bytes(crs), total, _u64(ksks), l0, count,
                                     _u64(zero), _u64(cts), out.reshape(-1))
        return out

    def ni_share_auto_ksk(self, crs, noise_seed, s, z):
        """One party's automorphism-key b parts [(w+1)][d_auto][N]
        (coefficient domain) and LWE-KSK b parts [N * d_ks] (mod q_ks)."""
        p, N = self.p, self.N
        auto_b = np.zeros((p.w + 1, p.auto_d, N), dtype=np.uint64)
        ksk_b = np.zeros(N * p.lwe_d, dtype=np.uint64)
        self.L.pzo_ni_share_auto_ksk(self.ev, seed_bytes(crs), seed_bytes(noise_seed), _i32(s),
                                     _i32(z), auto_b.reshape(-1), ksk_b)
        return auto_b, ksk_b

    def ni_aggregate_auto_ksk(self, auto_shares, ksk_shares):
        p, N = self.p, self.N
        a = np.ascontiguousarray(auto_shares, dtype=np.uint64)
        k = np.ascontiguousarray(ksk_shares, dtype=np.uint64)
        auto_b = np.zeros((p.w + 1, p.auto_d, N), dtype=np.uint64)
        ksk_b = np.zeros(N * p.lwe_d, dtype=np.uint64)
        self.L.pzo_ni_aggregate_auto_ksk(self.ev, a.shape[0], a.reshape(-1), k.reshape(-1),
                                         auto_b.reshape(-1), ksk_b)
        return auto_b, ksk_b

    def ni_expand_server_key(self, crs, rgsw, auto_b, ksk_b):
        """Evaluation-domain server key (with Shoup twins) from the seeded
        server key's parts: coefficient-domain RGSWs and summed b parts."""
        key = OracleKey(self.sizes, self.N)
        self.L.pzo_ni_expand_server_key(self.ev, seed_bytes(crs), _u64(rgsw), _u64(auto_b),
                                        _u64(ksk_b), C.byref(key.cstruct()))
        return key

    # ---- encryption ----
    def encrypt_bool(self, s, bit, seed):
        out = np.zeros(self.N + 1, dtype=np.uint64)
        m = self.L.pzo_true_el(self.ev) if bit else self.L.pzo_false_el(self.ev)
        self.L.pzo_encrypt_lwe(self.ev, seed_bytes(seed), np.ascontiguousarray(s, dtype=np.int32),
                               self.N, m, out)
        return out

    def decrypt_bool(self, s, lwe):
        m = self.L.pzo_decrypt_lwe(self.ev, np.ascontiguousarray(lwe, dtype=np.uint64),
                                   np.ascontiguousarray(s, dtype=np.int32), len(s))
        return int(self.L.pzo_decode_bool(self.ev, m))

    def decrypt_raw(self, s, lwe):
        return int(self.L.pzo_decrypt_lwe(self.ev, np.ascontiguousarray(lwe, dtype=np.uint64),
                                          np.ascontiguousarray(s, dtype=np.int32), len(s)))
