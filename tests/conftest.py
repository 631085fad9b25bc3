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
