"""
Compiled contraction programs (``include/tensorsat_prog.h``): jobs of pairwise steps
with every intermediate's axis order and arena offset fixed on the host.

A program is the compiled form of the contract2 argument lists that the reference's
``SimpleContraction._new`` precomputes (``contractions/simple.py:134-194``) and of the
pairwise loop cotengra runs (``contractions/cotengra.py:120-126``).  The same records
drive the GPU executors and the CPU checker.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import plan as _plan
from .network import Wiring
from .sat import CNFInstance

MAX_BITS = 27
NOPOS = 0xFF
STEP_KINNER = 1
STEP_NORHS = 2
STEP_PACKET = 4  # packet form (packetize): pad[0] = first lane of the step
STEP_PEND = 8  # packet form: last record of its packet
PACKET_LANES = 64
PACKET_CHUNK = 16
NSTAT = 10

STEP_DTYPE = np.dtype(
    [
        ("lhs_off", "<u4"), ("rhs_off", "<u4"), ("out_off", "<u4"),
        ("lhs_rank", "u1"), ("rhs_rank", "u1"), ("out_rank", "u1"), ("sum_rank", "u1"),
        ("flags", "u1"), ("pad", "u1", (3,)),
        ("o_lpos", "u1", (MAX_BITS,)), ("o_rpos", "u1", (MAX_BITS,)),
        ("c_lpos", "u1", (MAX_BITS,)), ("c_rpos", "u1", (MAX_BITS,)),
    ]
)
assert STEP_DTYPE.itemsize == 128

JOB_DTYPE = np.dtype(
    [
        ("instance", "<i4"), ("cube", "<i4"), ("step_begin", "<i4"), ("num_steps", "<i4"),
        ("leaf_begin", "<i8"), ("num_leaf_bytes", "<i4"), ("result_off", "<i4"),
        ("result_rank", "<i4"), ("free_vars", "<i4"), ("arena_elems", "<i8"),
        ("max_rank", "<i4"), ("flags", "<i4"),
    ]
)
assert JOB_DTYPE.itemsize == 56

STAT_NAMES = ("jobs", "branch_nodes", "pruned", "truncated", "max_rank", "host_verdict",
              "plan_us", "empty_residuals", "learned_trees", "fallback_trees")


class _SatOpts(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32), ("max_cubes", ctypes.c_int32), ("plan_trials", ctypes.c_int32),
        ("threads", ctypes.c_int32), ("seed", ctypes.c_uint64), ("cube_begin", ctypes.c_int32),
        ("priority", ctypes.c_void_p), ("gate_margin", ctypes.c_float), ("max_nodes", ctypes.c_int32),
        ("stop_on_sat", ctypes.c_int32), ("prior_branch", ctypes.c_int32), ("prior_lam", ctypes.c_float),
    ]


_bound = False


def _lib():
    global _bound
    L = _plan.lib()
    if not _bound:
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        L.tsc_compile.argtypes = [P(_plan._Net), P(ctypes.c_int32), P(ctypes.c_void_p), P(vp)]
        L.tsc_sat_batch.argtypes = [ctypes.c_int32, P(ctypes.c_int32), P(ctypes.c_int64), P(ctypes.c_int64),
                                    P(ctypes.c_int32), P(_SatOpts), P(vp)]
        L.tsc_program_sizes.argtypes = [vp, P(ctypes.c_int64), P(ctypes.c_int64), P(ctypes.c_int64), P(ctypes.c_int64)]
        L.tsc_program_copy.argtypes = [vp, vp, vp, vp, vp]
        L.tsc_program_free.argtypes = [vp]
        L.tsc_program_free.restype = None
        L.tsp_elim.argtypes = [P(_plan._Net), ctypes.c_int32, ctypes.c_uint64, P(ctypes.c_int32), P(_plan._Cost)]
        L.tsp_elim_order.argtypes = [P(_plan._Net), ctypes.c_int32, ctypes.c_uint64, P(ctypes.c_int32),
                                     P(_plan._Cost)]
        L.tsp_order_cost.argtypes = [P(_plan._Net), P(ctypes.c_int32), P(ctypes.c_int32), P(_plan._Cost)]
        L.tsp_elim_order_prior.argtypes = [P(_plan._Net), P(ctypes.c_float), ctypes.c_float, P(ctypes.c_int32),
                                           P(_plan._Cost)]
        L.tsc_packetize.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int64, ctypes.c_int32]
        L.tsc_search_pack.argtypes = [ctypes.c_int32, vp, vp, vp, vp, vp, vp, P(ctypes.c_int64), P(ctypes.c_int32),
                                      ctypes.c_int32]
        L.tsc_search_pack.restype = ctypes.c_int
        L.tsc_search_plan.argtypes = [ctypes.c_int32, vp, ctypes.c_double, ctypes.c_int32, ctypes.c_double,
                                      ctypes.c_int32, vp, ctypes.c_int64, P(ctypes.c_int64)]
        L.tsc_search_plan.restype = ctypes.c_int
        L.tsc_lanes.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int64, vp, ctypes.c_int64, vp, ctypes.c_int32]
        for f in (L.tsc_compile, L.tsc_sat_batch, L.tsc_program_sizes, L.tsc_program_copy, L.tsp_elim,
                  L.tsp_elim_order, L.tsp_order_cost, L.tsp_elim_order_prior, L.tsc_packetize, L.tsc_lanes):
            f.restype = ctypes.c_int
        _bound = True
    return L


@dataclass
class Program:
    """Jobs, steps and leaf bytes of a compiled program, plus per-instance statistics."""

    jobs: np.ndarray     # JOB_DTYPE
    steps: np.ndarray    # STEP_DTYPE
    leaves: np.ndarray   # uint8
    stats: np.ndarray    # int32 [num_instances, NSTAT]
    meta: dict = field(default_factory=dict)

    @property
    def num_instances(self) -> int:
        return int(self.stats.shape[0])

    def job_steps(self, j: int) -> np.ndarray:
        b = int(self.jobs["step_begin"][j])
        return self.steps[b : b + int(self.jobs["num_steps"][j])]

    def job_leaves(self, j: int) -> np.ndarray:
        b = int(self.jobs["leaf_begin"][j])
        return self.leaves[b : b + int(self.jobs["num_leaf_bytes"][j])]


def _take(handle) -> Program:
    L = _lib()
    nj, ns, nl, ni = (ctypes.c_int64() for _ in range(4))
    try:
        _plan._check(L.tsc_program_sizes(handle, ctypes.byref(nj), ctypes.byref(ns), ctypes.byref(nl),
                                         ctypes.byref(ni)), "tsc_program_sizes")
        jobs = np.empty(nj.value, dtype=JOB_DTYPE)  # every byte written by tsc_program_copy
        steps = np.empty(ns.value, dtype=STEP_DTYPE)
        leaves = np.empty(nl.value, dtype=np.uint8)
        stats = np.zeros((ni.value, NSTAT), dtype=np.int32)
        _plan._check(L.tsc_program_copy(handle, jobs.ctypes.data, steps.ctypes.data, leaves.ctypes.data,
                                        stats.ctypes.data), "tsc_program_copy")
    finally:
        L.tsc_program_free(handle)
    return Program(jobs, steps, leaves, stats)


def elim(wiring: Wiring, *, trials: int = 8, seed: int = 0) -> _plan.Plan:
    """Elimination-order planner (min-fill / min-degree trials)."""
    L = _lib()
    h = _plan._NetHolder(wiring)
    S = wiring.num_slots
    path = np.zeros(max(0, 2 * (S - 1)), dtype=np.int32)
    c = _plan._Cost()
    _plan._check(L.tsp_elim(ctypes.byref(h.net), trials, seed, _plan._ptr(path), ctypes.byref(c)), "tsp_elim")
    return _plan.Plan(S, path.reshape(-1, 2), cost=_plan.Cost._from(c))


def elim_order(wiring: Wiring, *, trials: int = 8, seed: int = 0) -> tuple[np.ndarray, _plan.Cost]:
    """The cheapest of ``trials`` min-fill/min-degree elimination orders and its tree cost."""
    L = _lib()
    h = _plan._NetHolder(wiring)
    order = np.zeros(max(1, wiring.num_wires), dtype=np.int32)
    c = _plan._Cost()
    _plan._check(L.tsp_elim_order(ctypes.byref(h.net), trials, seed, _plan._ptr(order), ctypes.byref(c)),
                 "tsp_elim_order")
    return order[: wiring.num_wires], _plan.Cost._from(c)


def order_cost(wiring: Wiring, order) -> tuple[_plan.Plan, _plan.Cost]:
    """Tree and exact cost of an explicit elimination order over all wires."""
    L = _lib()
    h = _plan._NetHolder(wiring)
    order = np.ascontiguousarray(np.asarray(order, dtype=np.int32))
    if order.shape != (wiring.num_wires,) or sorted(order.tolist()) != list(range(wiring.num_wires)):
        raise ValueError("order must be a permutation of the wires")
    S = wiring.num_slots
    path = np.zeros(max(1, 2 * (S - 1)), dtype=np.int32)
    c = _plan._Cost()
    _plan._check(L.tsp_order_cost(ctypes.byref(h.net), _plan._ptr(order), _plan._ptr(path), ctypes.byref(c)),
                 "tsp_order_cost")
    cost = _plan.Cost._from(c)
    return _plan.Plan(S, path[: max(0, 2 * (S - 1))].reshape(-1, 2), cost=cost), cost


def step_work(steps: np.ndarray, min_sum_bits: int = 6) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """
    Per step record: (takes the int8 MFMA route, operations 2^(out + summed + 1), bytes
    2^lhs + 2^rhs + 2^out).  The route rule is the executor's (``ts_dense_config``'s
    ``min_sum_bits``, 6 by default): a binary step whose summed bits are innermost in both
    operands (TS_STEP_KINNER), with at least ``min_sum_bits`` summed bits and at least 4
    out bits held by only one operand on each side.
    """
    out_rank = steps["out_rank"].astype(np.int64)
    sum_rank = steps["sum_rank"].astype(np.int64)
    binary = (steps["flags"] & 2) == 0
    kinner = (steps["flags"] & STEP_KINNER) != 0
    bit = np.arange(MAX_BITS)[None, :] < out_rank[:, None]
    has_l = steps["o_lpos"] != NOPOS
    has_r = steps["o_rpos"] != NOPOS
    mb = (bit & has_l & ~has_r).sum(axis=1)
    nb = (bit & has_r & ~has_l).sum(axis=1)
    dense = binary & kinner & (sum_rank >= min_sum_bits) & (mb >= 4) & (nb >= 4)
    ops = np.ldexp(1.0, out_rank + sum_rank + 1)
    rhs = np.where(binary, np.ldexp(1.0, steps["rhs_rank"].astype(np.int64)), 0.0)
    nbytes = np.ldexp(1.0, steps["lhs_rank"].astype(np.int64)) + rhs + np.ldexp(1.0, out_rank)
    return dense, ops, nbytes


def compile_network(wiring: Wiring, leaves, plan: _plan.Plan) -> Program:
    """
    Compile a bit network (every wire extent 2, unique wires per slot, unique out wires)
    with host leaf tensors (uint8, slot wire order) along ``plan``.
    """
    if any(d != 2 for d in wiring.wire_dims):
        raise ValueError("compile_network needs extent-2 wires")
    if len(leaves) != wiring.num_slots:
        raise ValueError("one leaf per slot is required")
    arrs = [np.ascontiguousarray(np.asarray(t, dtype=np.uint8)) for t in leaves]
    for s, (a, ws) in enumerate(zip(arrs, wiring.slot_wires_list)):
        if a.size != 1 << len(ws):
            raise ValueError(f"leaf {s} has {a.size} entries, expected {1 << len(ws)}")
    ptrs = (ctypes.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
    path = np.ascontiguousarray(plan.path.reshape(-1), dtype=np.int32)
    return _compile_ptrs(_plan._NetHolder(wiring), path, ptrs, plan)


def _compile_ptrs(h, path: np.ndarray, ptrs, plan: _plan.Plan) -> Program:
    """``tsc_compile`` of a network descriptor, a flat path and leaf pointers (the leaf
    bytes are copied during the call)."""
    handle = ctypes.c_void_p()
    _plan._check(_lib().tsc_compile(ctypes.byref(h.net), _plan._ptr(path), ptrs, ctypes.byref(handle)), "tsc_compile")
    prog = _take(handle)
    prog.meta["plan"] = plan
    return prog


def compile_sliced(wiring: Wiring, leaves, plan: _plan.Plan, sliced, threads: int = 0) -> list[Program]:
    """
    One compiled program per assignment of the ``sliced`` wires (in
    ``itertools.product(range(2), repeat=len(sliced))`` order): ``compile_network`` of the
    wiring without the sliced wires, with every leaf that holds one of them fixed at
    that assignment (``Box.__getitem__``).  The network, the path and the pointers of the
    leaves that hold no sliced wire are built once; per slice only the fixed leaves are
    cut and their pointers replaced.  Slices compile on ``threads`` threads (0: up to 8;
    the native compiler releases the interpreter lock and keeps its scratch per thread).
    """
    import itertools
    import os
    from concurrent.futures import ThreadPoolExecutor

    sliced = list(sliced)
    drop = set(sliced)
    full = wiring.slot_wires_list
    if any(d != 2 for d in wiring.wire_dims):
        raise ValueError("compile_sliced needs extent-2 wires")
    if len(leaves) != wiring.num_slots:
        raise ValueError("one leaf per slot is required")
    arrs = [np.ascontiguousarray(np.asarray(t, dtype=np.uint8)) for t in leaves]
    for s, (a, ws) in enumerate(zip(arrs, full)):
        if a.size != 1 << len(ws):
            raise ValueError(f"leaf {s} has {a.size} entries, expected {1 << len(ws)}")
    sub = Wiring(wiring.wire_dims, [tuple(x for x in ws if x not in drop) for ws in full], ())
    hot = [s for s, ws in enumerate(full) if drop.intersection(ws)]
    # Every cut of every leaf that holds sliced wires, once, in one buffer: leaf r's cut
    # for local assignment a (bit j = the value of its j-th sliced wire, in slot order)
    # starts at var_off[r] + a * var_size[r].
    sidx = {w: i for i, w in enumerate(sliced)}
    weights = np.zeros((len(hot), len(sliced)), dtype=np.int64)
    var_off, var_size, chunks, off = [], [], [], 0
    for r, s in enumerate(hot):
        ws = full[s]
        sw = [x for x in ws if x in drop]
        for j, x in enumerate(sw):
            weights[r, sidx[x]] = 1 << j
        shaped = arrs[s].reshape((2,) * len(ws))
        size = 1 << (len(ws) - len(sw))
        var_off.append(off)
        var_size.append(size)
        for a in range(1 << len(sw)):
            fixed = {x: (a >> j) & 1 for j, x in enumerate(sw)}
            chunks.append(np.ascontiguousarray(shaped[tuple(fixed.get(x, slice(None)) for x in ws)]).reshape(-1))
            off += size
    variants = np.concatenate(chunks) if chunks else np.zeros(1, dtype=np.uint8)
    hot_idx = np.asarray(hot, dtype=np.int64)
    hot_base = np.uint64(variants.ctypes.data) + np.asarray(var_off, dtype=np.uint64)
    hot_size = np.asarray(var_size, dtype=np.uint64)
    base = np.asarray([a.ctypes.data for a in arrs] or [0], dtype=np.uint64)
    PtrArray = ctypes.c_void_p * len(base)
    h = _plan._NetHolder(sub)
    path = np.ascontiguousarray(plan.path.reshape(-1), dtype=np.int32)

    def one(vals) -> Program:
        local = (weights @ np.asarray(vals, dtype=np.int64)).astype(np.uint64)
        p = base.copy()
        p[hot_idx] = hot_base + local * hot_size
        return _compile_ptrs(h, path, PtrArray.from_buffer_copy(p), plan)  # variants outlive the call

    assignments = list(itertools.product(range(2), repeat=len(sliced)))
    nt = threads if threads > 0 else min(8, os.cpu_count() or 1)
    if nt <= 1 or len(assignments) < 8:
        return [one(v) for v in assignments]
    with ThreadPoolExecutor(max_workers=nt) as ex:
        return list(ex.map(one, assignments))


def sat_batch(
    cnfs: list[CNFInstance],
    *,
    width: int = 18,
    max_cubes: int = 0,
    plan_trials: int = 2,
    threads: int = 0,
    seed: int = 0,
    cube_begin: int = 0,
    priority: np.ndarray | None = None,
    gate_margin: float = 0.0,
    max_nodes: int = 0,
    stop_on_sat: bool = False,
    prior_branch: bool = False,
    prior_lam: float = 16.0,
) -> Program:
    """
    Conditioned-slicing front-end for a batch of CNF instances: one job per cube, in
    depth-first order per instance (``tsc_sat_batch``).  ``cube_begin`` skips the first
    cubes of each instance (compiled by an earlier pass) and ``max_cubes`` caps the jobs
    compiled per instance from there; ``stats[:, 3]`` flags instances that were cut off.
    """
    nv = np.asarray([c.num_vars for c in cnfs], dtype=np.int32)
    clause_counts = np.asarray([len(c.clauses) for c in cnfs], dtype=np.int64)
    coff = np.zeros(len(cnfs) + 1, dtype=np.int64)
    coff[1:] = np.cumsum(clause_counts)
    lens = np.fromiter((len(cl) for c in cnfs for cl in c.clauses), dtype=np.int64, count=int(coff[-1]))
    loff = np.zeros(int(coff[-1]) + 1, dtype=np.int64)
    loff[1:] = np.cumsum(lens)
    lits = np.fromiter((l for c in cnfs for cl in c.clauses for l in cl), dtype=np.int32, count=int(loff[-1]))
    prio = None
    if priority is not None:
        prio = np.ascontiguousarray(np.asarray(priority, dtype=np.float32))
        if prio.shape != (int(nv.sum()),):
            raise ValueError("priority needs one value per variable of every instance")
    opts = _SatOpts(width, max_cubes, plan_trials, threads, seed, cube_begin,
                    prio.ctypes.data if prio is not None else None, gate_margin, max_nodes, int(stop_on_sat),
                    int(prior_branch), float(prior_lam))
    handle = ctypes.c_void_p()
    L = _lib()
    P = ctypes.POINTER
    rc = L.tsc_sat_batch(
        len(cnfs), nv.ctypes.data_as(P(ctypes.c_int32)), coff.ctypes.data_as(P(ctypes.c_int64)),
        loff.ctypes.data_as(P(ctypes.c_int64)), lits.ctypes.data_as(P(ctypes.c_int32)),
        ctypes.byref(opts), ctypes.byref(handle),
    )
    _plan._check(rc, "tsc_sat_batch")
    prog = _take(handle)
    prog.meta.update(width=width, max_cubes=max_cubes, plan_trials=plan_trials, cube_begin=cube_begin)
    return prog


def concat(progs: list[Program], instances: list[np.ndarray], num_instances: int) -> Program:
    """
    One program holding the jobs of ``progs``, program k's instance i renamed
    ``instances[k][i]`` (in ``range(num_instances)``), so that one persistent launch runs
    them all and early exit sees every cube of an instance.  Steps and leaves are
    appended (leaf blocks stay 16-byte aligned); ``stats`` rows are taken from the first
    program that holds each instance.
    """
    jobs, steps, leaves = [], [], []
    nsteps = nleaves = 0
    stats = np.zeros((num_instances, NSTAT), dtype=np.int32)
    have = np.zeros(num_instances, dtype=bool)
    for p, m in zip(progs, instances):
        m = np.asarray(m, dtype=np.int64)
        if m.shape != (p.num_instances,) or (m.size and (m.min() < 0 or m.max() >= num_instances)):
            raise ValueError("instance map must rename every instance into range(num_instances)")
        j = p.jobs.copy()
        j["step_begin"] += nsteps
        j["leaf_begin"] += nleaves
        j["instance"] = m[j["instance"]] if len(j) else j["instance"]
        jobs.append(j)
        steps.append(p.steps)
        pad = (-p.leaves.size) % 16
        leaves.append(np.concatenate([p.leaves, np.zeros(pad, np.uint8)]) if pad else p.leaves)
        nsteps += len(p.steps)
        nleaves += p.leaves.size + pad
        new = ~have[m]
        stats[m[new]] = p.stats[new]
        have[m] = True
    if nsteps >= 2**31:
        raise ValueError("too many steps for one program")
    out = Program(np.concatenate(jobs) if jobs else np.zeros(0, JOB_DTYPE),
                  np.concatenate(steps) if steps else np.zeros(0, STEP_DTYPE),
                  np.concatenate(leaves) if leaves else np.zeros(0, np.uint8), stats)
    out.meta["parts"] = [len(p.jobs) for p in progs]
    return out


def packetize(prog: Program, threads: int = 0) -> Program:
    """
    The packet form of ``prog`` for one-wave workgroups (``tsc_packetize``): each job's
    steps regrouped into packets of mutually independent steps sharing a wave's 64 lanes,
    intermediates reallocated over packet lifetimes.  Every step must have at most 6 out
    bits.  Values are unchanged (run in record order, the steps compute what the original
    program computes); returns a new Program.
    """
    jobs = prog.jobs.copy()
    steps = prog.steps.copy()
    if len(steps) and int(steps["out_rank"].max()) > 6:
        raise ValueError("packet form needs steps of at most 6 out bits")
    _plan._check(_lib().tsc_packetize(jobs.ctypes.data, len(jobs), steps.ctypes.data, len(steps), int(threads)),
                 "tsc_packetize")
    out = Program(jobs, steps, prog.leaves, prog.stats, dict(prog.meta))
    out.meta["packetized"] = True
    return out


LANE_DTYPE = np.dtype(
    [("lhs", "<u2"), ("rhs", "<u2"), ("out", "<u2"), ("tl", "<u2"), ("tr", "<u2"), ("op", "u1"), ("first", "u1"),
     ("count", "u1"), ("pad", "u1", (3,))]
)
assert LANE_DTYPE.itemsize == 16
LANE_NONE = 0xFFFF
LANE_GENERAL = 8
LANE_NORHS = 16
LANE_SKIP = 32
LANE_CHUNK = 128


@dataclass
class LaneProgram:
    """Lane form (``tsc_lanes``) of a packet-form program: ``jobs`` are the program's jobs
    with step_begin / num_steps counting lane records, ``lanes`` the records."""

    jobs: np.ndarray
    lanes: np.ndarray
    leaves: np.ndarray
    stats: np.ndarray
    source: Program

    @property
    def num_instances(self) -> int:
        return int(self.stats.shape[0])

    def job_lanes(self, j: int) -> np.ndarray:
        b = int(self.jobs["step_begin"][j])
        return self.lanes[b : b + int(self.jobs["num_steps"][j])]


def lanes(pk: Program, threads: int = 0) -> LaneProgram:
    """
    Lane form of a packet-form program: one 16-byte record per lane of every packet with
    absolute arena offsets (all index math on the host), packets padded so that none
    straddles a ``LANE_CHUNK``-record boundary.  Needs steps of at most 6 out and 6
    summed bits and arenas below 64 KiB.
    """
    L = _lib()
    nj = len(pk.jobs)
    begin = np.zeros(nj + 1, np.int64)
    jobs = np.ascontiguousarray(pk.jobs)
    steps = np.ascontiguousarray(pk.steps)
    _plan._check(L.tsc_lanes(jobs.ctypes.data, nj, steps.ctypes.data, len(steps), None, 0, begin.ctypes.data,
                             int(threads)), "tsc_lanes")
    total = int(begin[-1])
    if total >= 2**31:
        raise ValueError("too many lane records for one program")
    out = np.zeros(max(1, total), LANE_DTYPE)
    _plan._check(L.tsc_lanes(jobs.ctypes.data, nj, steps.ctypes.data, len(steps), out.ctypes.data, total,
                             begin.ctypes.data, int(threads)), "tsc_lanes")
    lj = jobs.copy()
    lj["step_begin"] = begin[:-1]
    lj["num_steps"] = np.diff(begin)
    return LaneProgram(lj, out[:total], pk.leaves, pk.stats, pk)


def validate_lanes(lp: LaneProgram) -> None:
    """Host-side bounds check of a lane-form program before any device launch."""
    jobs = lp.jobs
    if len(jobs) == 0:
        return
    if np.any(jobs["arena_elems"] > 0xFFFF):
        raise ValueError("lane form needs arenas below 64 KiB")
    if np.any(jobs["step_begin"].astype(np.int64) + jobs["num_steps"] > len(lp.lanes)) or np.any(jobs["step_begin"] < 0):
        raise ValueError("lane range outside the lane array")
    if np.any(jobs["leaf_begin"] + jobs["num_leaf_bytes"] > lp.leaves.size) or np.any(jobs["leaf_begin"] % 16):
        raise ValueError("leaf range outside the leaf buffer or unaligned")
    if np.any(jobs["num_leaf_bytes"].astype(np.int64) > jobs["arena_elems"]):
        raise ValueError("leaves exceed the job arena")
    if np.any(jobs["result_off"].astype(np.int64) + (1 << jobs["result_rank"].astype(np.int64)) > jobs["arena_elems"]):
        raise ValueError("result outside the job arena")
    if len(lp.lanes) == 0:
        return
    owner = np.repeat(np.arange(len(jobs)), jobs["num_steps"])
    ln = lp.lanes[np.concatenate([np.arange(b, b + n) for b, n in zip(jobs["step_begin"], jobs["num_steps"])])]
    arena = jobs["arena_elems"][owner].astype(np.int64)
    nc = (ln["op"] & 7).astype(np.int64)
    skip = (ln["op"] & LANE_SKIP) != 0
    general = (ln["op"] & LANE_GENERAL) != 0
    # K-innermost / sum-free lanes read 2^nc contiguous bytes; a general lane's terms add
    # table offsets whose bits are disjoint from its own (in bounds by construction).
    span = np.where(general, 1, 1 << nc)
    ok = skip | ((ln["lhs"].astype(np.int64) + span <= arena) & (ln["rhs"].astype(np.int64) + span <= arena))
    ok &= skip | ~general | ((ln["tl"].astype(np.int64) < arena) & (ln["tr"].astype(np.int64) < arena))
    ok &= skip | (ln["out"] == LANE_NONE) | (ln["out"].astype(np.int64) < arena)
    if not np.all(ok):
        raise ValueError("lane record outside its job arena")


def packet_spans(steps: np.ndarray) -> list[tuple[int, int]]:
    """[begin, end) record ranges of the packets of one packetized job's steps."""
    ends = np.flatnonzero(steps["flags"] & STEP_PEND) + 1
    begins = np.concatenate([[0], ends[:-1]]) if len(ends) else np.zeros(0, np.int64)
    return [(int(b), int(e)) for b, e in zip(begins, ends)]


def step_labels(st) -> tuple[list[int], list[int], list[int]]:
    """
    Einsum labels of one step record: out bit j is label j, summed bit j is label
    ``out_rank + j``; operand axes are listed MSB first (row-major order).
    """
    ro, nc = int(st["out_rank"]), int(st["sum_rank"])
    rl, rr = int(st["lhs_rank"]), int(st["rhs_rank"])
    lab_l = [-1] * rl
    lab_r = [-1] * rr
    for j in range(ro):
        p = int(st["o_lpos"][j])
        if p != NOPOS:
            lab_l[rl - 1 - p] = j
        p = int(st["o_rpos"][j])
        if p != NOPOS:
            lab_r[rr - 1 - p] = j
    for j in range(nc):
        p = int(st["c_lpos"][j])
        if p != NOPOS:
            lab_l[rl - 1 - p] = ro + j
        p = int(st["c_rpos"][j])
        if p != NOPOS:
            lab_r[rr - 1 - p] = ro + j
    out = [ro - 1 - i for i in range(ro)]
    if -1 in lab_l or -1 in lab_r:
        raise ValueError("malformed step record")
    return lab_l, lab_r, out


def validate(prog: Program) -> None:
    """
    Host-side bounds check of every record before any device launch: operand and out
    ranges inside the job arena, ranks within limits, leaves inside the arena and the
    leaf buffer, and summed/out bit positions inside each operand.
    """
    jobs, steps = prog.jobs, prog.steps
    if len(jobs) == 0:
        return
    if np.any(jobs["num_leaf_bytes"].astype(np.int64) > jobs["arena_elems"]):
        raise ValueError("leaves exceed the job arena")
    if np.any(jobs["leaf_begin"] + jobs["num_leaf_bytes"] > prog.leaves.size):
        raise ValueError("leaf range outside the leaf buffer")
    if np.any(jobs["leaf_begin"] % 16):
        raise ValueError("leaf blocks must be 16-byte aligned")
    if np.any(jobs["step_begin"] + jobs["num_steps"] > len(steps)) or np.any(jobs["step_begin"] < 0):
        raise ValueError("step range outside the step array")
    if np.any(jobs["result_off"].astype(np.int64) + (1 << jobs["result_rank"].astype(np.int64)) > jobs["arena_elems"]):
        raise ValueError("result outside the job arena")
    if len(steps) == 0:
        return
    owner = np.repeat(np.arange(len(jobs)), jobs["num_steps"])
    order = np.concatenate([np.arange(b, b + n) for b, n in zip(jobs["step_begin"], jobs["num_steps"])]) \
        if len(owner) else np.zeros(0, dtype=np.int64)
    st = steps[order]
    arena = jobs["arena_elems"][owner].astype(np.int64)
    for name in ("lhs_rank", "rhs_rank", "out_rank", "sum_rank"):
        if np.any(st[name] > MAX_BITS):
            raise ValueError(f"{name} above {MAX_BITS}")
    norhs = (st["flags"] & STEP_NORHS) != 0
    one = np.int64(1)
    if np.any(st["lhs_off"].astype(np.int64) + (one << st["lhs_rank"].astype(np.int64)) > arena):
        raise ValueError("lhs outside the job arena")
    if np.any(~norhs & (st["rhs_off"].astype(np.int64) + (one << st["rhs_rank"].astype(np.int64)) > arena)):
        raise ValueError("rhs outside the job arena")
    if np.any(st["out_off"].astype(np.int64) + (one << st["out_rank"].astype(np.int64)) > arena):
        raise ValueError("out outside the job arena")
    packet = (st["flags"] & STEP_PACKET) != 0
    if np.any(packet & (st["out_rank"] > 6)):
        raise ValueError("packet-form steps have at most 6 out bits")
    for pos, rank in (("o_lpos", "lhs_rank"), ("c_lpos", "lhs_rank"), ("o_rpos", "rhs_rank"), ("c_rpos", "rhs_rank")):
        p = st[pos].astype(np.int64)
        lim = st[rank].astype(np.int64)[:, None]
        valid = (p == NOPOS) | (p < lim)
        if pos == "o_lpos":  # packet form: bytes 16..23 carry the packet's lane mask
            valid[:, 16:24] |= packet[:, None]
        if not np.all(valid):
            raise ValueError(f"{pos} outside its operand")
    if np.any(st["lhs_off"] % 16) or np.any(~norhs & (st["rhs_off"] % 16 != 0)) or np.any(st["out_off"] % 16):
        raise ValueError("tensor offsets must be 16-element aligned")
