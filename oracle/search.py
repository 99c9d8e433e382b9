"""
ORACLE — TEST INFRASTRUCTURE ONLY.  ctypes wrapper of ``oracle/search.c``: the CPU
restatement of the device cube search that decides batched instances (conditioned
slicing by a depth-first branch, residuals of at most ``width`` variables contracted
into one relation and reduced).  Built by ``make -C oracle`` into ``oracle/_build``.
It is also timed, on all host cores, beside the device in ``bench.py``.

Pinning: ``tests/test_search_oracle.py`` (verdicts against brute force and DPLL; leaf
relations against :mod:`oracle.finrel`).  Branch statistics: parity unpinned beyond
those (no reference counterpart).
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_LIB = _DIR / "_build" / "libsearch.so"
_lib = None

STAT_DTYPE = np.dtype([("verdict", "<i4"), ("max_depth", "<i4"), ("nodes", "<i8"), ("leaves", "<i8"),
                       ("conflicts", "<i8"), ("digest", "<u8")])
LEAF_DTYPE = np.dtype([("instance", "<i4"), ("rank", "<i4"), ("bits", "<u8"), ("vars", "<i4", (6,)),
                       ("pad", "<i4", (2,)), ("assigned", "<u8", (2,)), ("value", "<u8", (2,))])


def _load():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            subprocess.run(["make", "-C", str(_DIR)], check=True, capture_output=True)
        L = ctypes.CDLL(str(_LIB))
        vp = ctypes.c_void_p
        L.search_batch.argtypes = [ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int64, ctypes.c_int, vp, vp, vp, ctypes.c_int, vp]
        L.search_batch.restype = ctypes.c_int
        L.search_leaf_mix.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.search_leaf_mix.restype = ctypes.c_uint64
        _lib = L
    return _lib


def pack(cnfs):
    """(nvars, clause offsets per instance, literal offsets per clause, literals)."""
    nv = np.asarray([c.num_vars for c in cnfs], dtype=np.int32)
    coff = np.zeros(len(cnfs) + 1, dtype=np.int64)
    coff[1:] = np.cumsum([len(c.clauses) for c in cnfs])
    lens = np.fromiter((len(cl) for c in cnfs for cl in c.clauses), dtype=np.int64, count=int(coff[-1]))
    loff = np.zeros(int(coff[-1]) + 1, dtype=np.int64)
    loff[1:] = np.cumsum(lens)
    lits = np.fromiter((l for c in cnfs for cl in c.clauses for l in cl), dtype=np.int32, count=int(loff[-1]))
    return nv, coff, loff, lits


def search(cnfs, *, width: int = 6, exhaustive: bool = False, max_nodes: int = 1 << 40, threads: int = 1,
           trace: int = 0, packed=None, order: str = "score", witness: bool = False):
    """Per-instance statistics (STAT_DTYPE); with ``trace``, also the first leaves; with
    ``witness``, also each instance's first satisfiable leaf's assignment (a string of
    "0"/"1" per variable, None if none) -- with ``order="lex"`` the lexicographically
    smallest satisfying assignment."""
    nv, coff, loff, lits = packed if packed is not None else pack(cnfs)
    num = len(nv)
    out = np.zeros(num, dtype=STAT_DTYPE)
    tr = np.zeros(max(trace, 1), dtype=LEAF_DTYPE)
    wit = np.zeros((max(num, 1), 2), dtype=np.uint64)
    tl = ctypes.c_int(0)
    lp = lits if lits.size else np.zeros(1, np.int32)
    if order not in ("score", "lex"):
        raise ValueError("order must be 'score' or 'lex'")
    rc = _load().search_batch(num, nv.ctypes.data, coff.ctypes.data, loff.ctypes.data, lp.ctypes.data, width,
                              int(exhaustive), 1 if order == "lex" else 0, max_nodes, threads, out.ctypes.data,
                              wit.ctypes.data, tr.ctypes.data if trace else None, trace, ctypes.byref(tl))
    if rc == -2:
        raise ValueError("invalid instance (literal out of range, or width outside 0..6)")
    if rc == -3:
        raise ValueError("a clause has more than 3 distinct literals")
    res = [out]
    if trace:
        res.append(tr[: tl.value])
    if witness:
        ws = []
        for i in range(num):
            if out["verdict"][i] != 1:
                ws.append(None)
                continue
            bits = int(wit[i, 0]) | (int(wit[i, 1]) << 64)
            ws.append("".join("1" if (bits >> j) & 1 else "0" for j in range(int(nv[i]))))
        res.append(ws)
    return res[0] if len(res) == 1 else tuple(res)


def leaf_mix(bits: int, rank: int) -> int:
    return int(_load().search_leaf_mix(bits, rank))
