"""
ORACLE — TEST INFRASTRUCTURE ONLY.  Brute-force SAT / #SAT by enumeration.

Vectorised over all 2^n assignments (n <= 26), independent of any tensor-network
machinery: a clause is satisfied when one of its literals is true.  Pins verdicts and
model counts of small instances (``examples/test_sat_solver.py`` expectations,
``2025-02-13-#sat-tensor-contraction.ipynb`` cell 12 count).
"""

from __future__ import annotations

import numpy as np


def satisfying_mask(num_vars: int, clauses, chunk_bits: int = 22) -> np.ndarray:
    if num_vars > 26:
        raise ValueError("brute force limited to 26 variables")
    total = 1 << num_vars
    out = np.empty(total, dtype=bool)
    step = 1 << min(chunk_bits, num_vars)
    for start in range(0, total, step):
        idx = np.arange(start, min(total, start + step), dtype=np.int64)
        ok = np.ones(idx.size, dtype=bool)
        for cl in clauses:
            sat = np.zeros(idx.size, dtype=bool)
            for lit in cl:
                v = abs(lit) - 1
                # variable v is bit (n-1-v) of the assignment index (MSB = variable 0)
                bit = (idx >> (num_vars - 1 - v)) & 1
                sat |= bit.astype(bool) if lit > 0 else ~bit.astype(bool)
            ok &= sat
        out[start : start + idx.size] = ok
    return out


def is_sat(num_vars: int, clauses) -> bool:
    return bool(satisfying_mask(num_vars, clauses).any())


def count(num_vars: int, clauses) -> int:
    return int(satisfying_mask(num_vars, clauses).sum())


def solutions(num_vars: int, clauses) -> set[tuple[int, ...]]:
    m = satisfying_mask(num_vars, clauses)
    res = set()
    for i in np.flatnonzero(m):
        res.add(tuple((int(i) >> (num_vars - 1 - v)) & 1 for v in range(num_vars)))
    return res
