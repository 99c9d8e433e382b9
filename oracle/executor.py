"""
ORACLE — TEST INFRASTRUCTURE ONLY.  Reference-equivalent step loop over a program.

Runs the jobs of a compiled program (``tensorsat_amd.program``) on the CPU with numpy,
one pairwise step at a time, exactly as the reference's executors do: the contraction
loop of ``SimpleContraction.contract`` (``contractions/simple.py:251-292``) / cotengra's
``ContractionTree.contract`` (``contractions/cotengra.py:120-126``), each step being
``FinRel._contract2`` (``lang/fin_rel.py:255-288``) restated by :mod:`oracle.finrel`.

Semirings:

* ``"MOD256"`` — the reference's arithmetic: dtype widening by the contracted size
  (``fin_rel.py:145-159``), wrapped integer einsum, ``np.sign(..., dtype=uint8)``.
* ``"GT0"``    — ``count > 0`` with exact int64 counts.
* ``"COUNT"``  — exact counting; uint64 einsum when a bound shows it cannot wrap,
  Python-int object arrays otherwise; raises ``OverflowError`` if a value >= 2^64.

This is also the timed CPU baseline (``bench.py``): no autoray/cotengra/typing
overhead, so it is faster than the reference itself.
"""

from __future__ import annotations

from math import prod

import numpy as np

from . import finrel, records

# Decoded records of the programs seen last (keyed by the identity of their buffers).
_decoded: dict = {}


def _records(prog):
    """(jobs, steps) of a program, decoded from their raw bytes by :mod:`oracle.records`
    (the checker does not use the product's decoder)."""
    key = id(prog)
    hit = _decoded.get(key)
    if hit is not None and hit[0] is prog.steps and hit[1] is prog.jobs:
        return hit[2], hit[3]
    jobs = records.decode_jobs(prog.jobs)
    steps = records.decode_steps(prog.steps)
    if len(_decoded) > 8:
        _decoded.clear()
    _decoded[key] = (prog.steps, prog.jobs, jobs, steps)
    return jobs, steps


def _step(arena, st: records.Step, mode: str):
    lab_l, lab_r, lab_o = list(st.lhs_labels), list(st.rhs_labels), list(st.out_labels)
    rl, rr, ro = st.lhs_rank, st.rhs_rank, st.out_rank
    lo, ro_off, oo = st.lhs_off, st.rhs_off, st.out_off
    lhs = arena[lo : lo + (1 << rl)].reshape((2,) * rl)
    if st.norhs:
        rhs = np.ones((), dtype=arena.dtype)
        lab_r = []
    else:
        rhs = arena[ro_off : ro_off + (1 << rr)].reshape((2,) * rr)
    if mode == "GT0":
        cnt = np.einsum(lhs.astype(np.int64), lab_l, rhs.astype(np.int64), lab_r, lab_o)
        res = (cnt > 0).astype(np.uint8)
    elif mode == "MOD256":
        common = set(lab_l) & set(lab_r)
        csize = prod(2 for lab in common if lab not in lab_o)
        lt, rt = finrel.safe_contraction_dtype(csize, lhs, rhs)
        res = np.sign(np.einsum(lt, lab_l, rt, lab_r, lab_o), dtype=np.uint8)
    elif mode == "COUNT":
        nsum = st.sum_rank
        lmax = int(lhs.max()) if lhs.size else 0
        rmax = int(rhs.max()) if rhs.size else 0
        if lmax * rmax * (1 << nsum) < 2**64:
            res = np.einsum(lhs.astype(np.uint64), lab_l, rhs.astype(np.uint64), lab_r, lab_o)
        else:
            big = np.einsum(lhs.astype(object), lab_l, rhs.astype(object), lab_r, lab_o)
            if np.any(np.asarray(big, dtype=object) >= 2**64):
                raise OverflowError("count exceeds 2^64")
            res = np.asarray(big, dtype=np.uint64)
    else:
        raise ValueError(f"unknown mode {mode!r}")
    arena[oo : oo + (1 << ro)] = np.asarray(res).reshape(-1)
    return arena[oo : oo + (1 << ro)]


def run_job(prog, j: int, mode: str = "GT0", *, keep_intermediates: bool = False):
    """Run job ``j`` of a compiled program (its ``jobs`` / ``steps`` record arrays and
    ``leaves`` bytes); returns (result array, list of step outputs if requested)."""
    jobs, steps = _records(prog)
    job = jobs[j]
    dtype = np.uint64 if mode == "COUNT" else np.uint8
    arena = np.zeros(job.arena_elems, dtype=dtype)
    leaves = np.asarray(prog.leaves)[job.leaf_begin : job.leaf_begin + job.num_leaf_bytes]
    arena[: leaves.size] = leaves
    inter = []
    for st in steps[job.step_begin : job.step_begin + job.num_steps]:
        out = _step(arena, st, mode)
        if keep_intermediates:
            inter.append(out.copy())
    r0, rr = job.result_off, job.result_rank
    return arena[r0 : r0 + (1 << rr)].copy(), inter


def decide(prog, mode: str = "GT0") -> np.ndarray:
    """Per-instance verdicts: OR over each instance's job results (scalars)."""
    jobs, _ = _records(prog)
    out = np.zeros(prog.num_instances, dtype=np.uint8)
    for j, job in enumerate(jobs):
        if out[job.instance]:
            continue
        res, _ = run_job(prog, j, mode)
        out[job.instance] |= np.uint8(res.reshape(-1)[0] != 0)
    return out


def count(prog) -> list[int]:
    """Per-instance exact model counts: sum over jobs of result * 2^free_vars."""
    jobs, _ = _records(prog)
    out = [0] * prog.num_instances
    for j, job in enumerate(jobs):
        res, _ = run_job(prog, j, "COUNT")
        out[job.instance] += int(res.reshape(-1)[0]) << job.free_vars
    return out
