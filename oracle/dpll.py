"""
ORACLE — TEST INFRASTRUCTURE ONLY.  ctypes wrapper of ``oracle/dpll.c`` (an independent
DPLL decision procedure, the cross-check that kissat would provide).  Built by
``make -C oracle`` into ``oracle/_build``.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_LIB = _DIR / "_build" / "libdpll.so"
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            subprocess.run(["make", "-C", str(_DIR)], check=True, capture_output=True)
        L = ctypes.CDLL(str(_LIB))
        L.dpll_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
        L.dpll_solve.restype = ctypes.c_int
        _lib = L
    return _lib


def solve(num_vars: int, clauses, budget: int = 50_000_000) -> bool | None:
    """True SAT, False UNSAT, None if the node budget ran out."""
    off = np.zeros(len(clauses) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(c) for c in clauses])
    lits = np.fromiter((l for c in clauses for l in c), dtype=np.int32, count=int(off[-1]))
    if lits.size == 0:
        lits = np.zeros(1, np.int32)
    r = _load().dpll_solve(num_vars, len(clauses), off.ctypes.data, lits.ctypes.data, budget)
    if r == -2:
        raise ValueError("invalid clause list")
    return None if r == -1 else bool(r)
