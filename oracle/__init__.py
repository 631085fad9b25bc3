"""Python bindings of the C oracle (oracle/pz_oracle.c).

ORACLE / TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg, never by the product.
"""
from .oracle import Oracle, Params, build, PARAM_SETS, GATES  # noqa: F401
