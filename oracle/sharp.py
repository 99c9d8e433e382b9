"""
ORACLE — TEST INFRASTRUCTURE ONLY.  ctypes wrapper of ``oracle/sharp.c``: an exact
model counter (DPLL with connected components and a component cache) that shares no
code or algorithm with the product's tensor-contraction counts.  Built by
``make -C oracle`` into ``oracle/_build``.

Pinned by ``tests/test_oracle.py``: brute-force enumeration on random formulas of up to
16 variables (with unit, duplicate-literal, tautological and empty clauses) and the
#SAT notebook's count 6.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_LIB = _DIR / "_build" / "libsharp.so"
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            subprocess.run(["make", "-C", str(_DIR)], check=True, capture_output=True)
        L = ctypes.CDLL(str(_LIB))
        L.sharp_count.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                                  ctypes.c_void_p]
        L.sharp_count.restype = ctypes.c_int
        _lib = L
    return _lib


def count(num_vars: int, clauses, *, with_nodes: bool = False):
    """Exact number of models (Python int) of the CNF over ``num_vars`` variables."""
    off = np.zeros(len(clauses) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(c) for c in clauses])
    lits = np.fromiter((l for c in clauses for l in c), dtype=np.int32, count=int(off[-1]))
    if lits.size == 0:
        lits = np.zeros(1, np.int32)
    out = ctypes.create_string_buffer(64)
    nodes = ctypes.c_int64(0)
    rc = _load().sharp_count(num_vars, len(clauses), off.ctypes.data, lits.ctypes.data, out, ctypes.byref(nodes))
    if rc == -2:
        raise ValueError("invalid clause list")
    if rc == -3:
        raise OverflowError("model count reaches 2^128")
    c = int(out.value.decode())
    return (c, int(nodes.value)) if with_nodes else c
