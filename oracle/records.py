"""
ORACLE — TEST INFRASTRUCTURE ONLY.  Decoder of compiled-program records.

Reads the raw bytes of the job and step records from the byte layout documented in
``include/tensorsat_prog.h`` (``ts_job``, 56 bytes; ``ts_step``, 128 bytes), with its
own field offsets and its own einsum labelling, so that the checker shares no decoding
code with the product: a field misread by the product's ``program`` module would make
the two disagree instead of agreeing on the same mistake.

Step semantics (``tensorsat_prog.h``): for every out index ``o`` and summed index ``c``,
``acc(o) = sum_c lhs[L(o, c)] * rhs[R(o, c)]`` where out bit ``j`` of ``o`` sits at bit
``o_lpos[j]`` of the lhs flat index (``o_rpos[j]`` of the rhs; 0xFF = not held) and
summed bit ``j`` at ``c_lpos[j]`` / ``c_rpos[j]``; flat index bit ``b`` of a rank-r
tensor is its axis ``r - 1 - b`` (row-major).  That is ``FinRel._contract2``'s einsum
(``lang/fin_rel.py:255-288``) with the operands' wires spelled as bit positions.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

STEP_BYTES = 128
JOB_BYTES = 56
MAX_BITS = 27
ABSENT = 0xFF
FLAG_NORHS = 2

# ts_step: three u32 offsets, four u8 ranks, u8 flags, 3 pad bytes, then four position
# arrays of MAX_BITS bytes each.
_STEP_HEAD = struct.Struct("<3I4BB3x")
_POS0 = _STEP_HEAD.size  # 20
# ts_job: i32 instance, cube, step_begin, num_steps; i64 leaf_begin; i32 num_leaf_bytes,
# result_off, result_rank, free_vars; i64 arena_elems; i32 max_rank, flags.
_JOB = struct.Struct("<4iq4iq2i")
assert _POS0 + 4 * MAX_BITS == STEP_BYTES and _JOB.size == JOB_BYTES


@dataclass(frozen=True)
class Step:
    lhs_off: int
    rhs_off: int
    out_off: int
    lhs_rank: int
    rhs_rank: int
    out_rank: int
    sum_rank: int
    norhs: bool
    lhs_labels: tuple  # einsum labels of the lhs axes (axis 0 first)
    rhs_labels: tuple
    out_labels: tuple


@dataclass(frozen=True)
class Job:
    instance: int
    cube: int
    step_begin: int
    num_steps: int
    leaf_begin: int
    num_leaf_bytes: int
    result_off: int
    result_rank: int
    free_vars: int
    arena_elems: int


def _axis_labels(rank: int, pos_out, pos_sum, out_rank: int) -> tuple:
    """Label of each axis of an operand: out bit j -> label j, summed bit j -> label
    out_rank + j, placed at axis rank-1-position."""
    labels = [None] * rank
    for j, p in enumerate(pos_out):
        if p != ABSENT:
            if p >= rank or labels[rank - 1 - p] is not None:
                raise ValueError("step record places two bits on one axis or outside the operand")
            labels[rank - 1 - p] = j
    for j, p in enumerate(pos_sum):
        if p != ABSENT:
            if p >= rank or labels[rank - 1 - p] is not None:
                raise ValueError("step record places two bits on one axis or outside the operand")
            labels[rank - 1 - p] = out_rank + j
    if any(lab is None for lab in labels):
        raise ValueError("step record leaves an operand axis unlabelled")
    return tuple(labels)


def decode_step(raw: bytes) -> Step:
    lo, ro_, oo, rl, rr, rout, nsum, flags = _STEP_HEAD.unpack_from(raw, 0)
    if max(rout, nsum) > MAX_BITS:
        raise ValueError("step rank beyond the record's bit arrays")
    o_l = raw[_POS0 : _POS0 + rout]
    o_r = raw[_POS0 + MAX_BITS : _POS0 + MAX_BITS + rout]
    c_l = raw[_POS0 + 2 * MAX_BITS : _POS0 + 2 * MAX_BITS + nsum]
    c_r = raw[_POS0 + 3 * MAX_BITS : _POS0 + 3 * MAX_BITS + nsum]
    norhs = bool(flags & FLAG_NORHS)
    lab_l = _axis_labels(rl, o_l, c_l, rout)
    lab_r = () if norhs else _axis_labels(rr, o_r, c_r, rout)
    # out index bit j is out axis rout-1-j
    lab_o = tuple(rout - 1 - i for i in range(rout))
    return Step(lo, ro_, oo, rl, rr, rout, nsum, norhs, lab_l, lab_r, lab_o)


def decode_steps(buf) -> list[Step]:
    raw = bytes(memoryview(np.ascontiguousarray(buf)).cast("B"))
    if len(raw) % STEP_BYTES:
        raise ValueError("step buffer is not a whole number of records")
    return [decode_step(raw[i : i + STEP_BYTES]) for i in range(0, len(raw), STEP_BYTES)]


def decode_jobs(buf) -> list[Job]:
    raw = bytes(memoryview(np.ascontiguousarray(buf)).cast("B"))
    if len(raw) % JOB_BYTES:
        raise ValueError("job buffer is not a whole number of records")
    out = []
    for i in range(0, len(raw), JOB_BYTES):
        f = _JOB.unpack_from(raw, i)
        out.append(Job(f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], f[9]))
    return out
