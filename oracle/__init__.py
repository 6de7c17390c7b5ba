"""ORACLE -- test infrastructure only (the checker, never the product).

ctypes bindings for oracle/liboracle.so (the C restatement of ATLAS's CPU
path) and oracle/_ref/libatlref.so (ATLAS's own C reference BLAS, compiled
from /root/reference where that exists).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this package.
"""
from .orc import *  # noqa: F401,F403
