"""Shared pytest set-up: the `gpu` marker, import paths and cached fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "phantom-zone_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpz.so)")


_oracles = {}
_keys = {}


def oracle_for(name):
    from oracle import Oracle
    if name not in _oracles:
        _oracles[name] = Oracle(name)
    return _oracles[name]


# parties each parameter set is designed for (SP_TEST is the reference's
# single-party test set, src/bool/parameters.rs:706-726)
PARTIES = {"SP_TEST": 1, "NI_2P": 2, "NI_4P_HB_FR": 4, "NI_4P_LB_SR": 4, "NI_8P": 8,
           "I_2P_LB_SR": 2, "I_4P": 4, "I_8P": 8}


def keys_for(name, seed_secret=0, seed_key=1, parties=None):
    """(oracle, s, z, key) for a parameter set, generated once per session."""
    if parties is None:
        parties = PARTIES[name]
    k = (name, seed_secret, seed_key, parties)
    if k not in _keys:
        o = oracle_for(name)
        s, z, key = o.keygen(seed_secret, seed_key, parties)
        _keys[k] = (o, s, z, key)
    return _keys[k]


@pytest.fixture(scope="session")
def ni2p():
    return keys_for("NI_2P")


@pytest.fixture(scope="session")
def sptest():
    return keys_for("SP_TEST")


def rng(seed):
    return np.random.default_rng(seed)


def uniform_mod(r, q, shape):
    """uint64 values uniform in [0, q)."""
    return r.integers(0, q, size=shape, dtype=np.uint64)


def oracle_cone(o, key, nodes, inputs_ct, wires, threads=None):
    """Oracle ciphertexts of `wires` in a gate DAG (nodes from
    pzlib.Circuit.nodes(): kind 0 input, 1 gate, 2 NOT), evaluating only the
    wires' ancestors.  Every gate is a deterministic function of its two
    operands, so the words equal serial gate-by-gate evaluation in node order;
    the independent gates of one depth run on host threads."""
    nodes = [tuple(int(x) for x in nd) for nd in nodes]
    need, stack = set(), [int(w) for w in wires]
    while stack:
        i = stack.pop()
        if i in need:
            continue
        need.add(i)
        kind, _, a, b = nodes[i]
        if kind == 1:
            stack += [a, b]
        elif kind == 2:
            stack.append(a)
    vals, depth, k = {}, {}, 0
    for i, (kind, _, a, b) in enumerate(nodes):
        if kind == 0:
            if i in need:
                vals[i] = np.asarray(inputs_ct[k], dtype=np.uint64)
            depth[i] = 0
            k += 1
        elif kind == 1:
            depth[i] = max(depth[a], depth[b]) + 1
        else:
            depth[i] = depth[a]
    by_depth = {}
    for i in sorted(need):
        if nodes[i][0] != 0:
            by_depth.setdefault(depth[i], []).append(i)
    for d in sorted(by_depth):
        gates = [i for i in by_depth[d] if nodes[i][0] == 1]
        if gates:
            out = o.gates_batch(key, [nodes[i][1] for i in gates],
                                np.stack([vals[nodes[i][2]] for i in gates]),
                                np.stack([vals[nodes[i][3]] for i in gates]), threads)
            for j, i in enumerate(gates):
                vals[i] = out[j]
        for i in by_depth[d]:  # NOTs of this depth, in node order (they may chain)
            if nodes[i][0] == 2:
                vals[i] = o.not_(vals[nodes[i][2]])
    return {int(w): vals[int(w)] for w in wires}


def oracle_key_from_rows(o, rgsw, auto, ksk):
    """An oracle key holding server-key rows in the pz_key_upload layout (as
    pzlib's key_download returns them), with its Shoup twins."""
    import ctypes as C
    from oracle.oracle import OracleKey
    k = OracleKey(o.sizes, o.N)
    k.rgsw[:] = np.ascontiguousarray(rgsw, dtype=np.uint64).reshape(-1)
    k.auto_key[:] = np.ascontiguousarray(auto, dtype=np.uint64).reshape(-1)
    k.lwe_ksk[:] = np.ascontiguousarray(ksk, dtype=np.uint64).reshape(-1)
    o.L.pzo_key_shoup(o.ev, C.byref(k.cstruct()))
    return k
