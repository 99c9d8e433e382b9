"""
One large instance on the whole GPU.

* :func:`decide` — a SAT decision for one instance whose intermediates need the whole
  device (the n=120, m=510 workload: cubes of width up to 26, i.e. 2^26-entry tensors):
  cubes in depth-first order, in gangs (of 8 cubes, or 16 then 32 slices of one sliced
  tree) that advance together phase by phase: the runs of small steps of all the gang's cubes in one launch (one
  workgroup per cube), their large steps in one grid-wide launch; the flags are read
  after each gang and the first satisfiable cube stops the submission.
* :func:`decide_budget` — a SAT decision under a wall-clock budget for instances whose
  cube enumeration cannot be completed (the learned-planner workload, n = 150..300):
  cubes are compiled in chunks of doubling size in depth-first order and each chunk is
  decided on the GPU by the batched executor with early exit; the first satisfiable cube
  (or a residual with every clause satisfied, found on the host) decides SAT, a completed
  enumeration without one decides UNSAT, and otherwise the answer is unknown (-1).
* :func:`decide_learned` — the learned planner's search first, with the plain planner as
  fallback, in alternating time slices of doubling length under one budget.
* :func:`count` — the exact model count (#SAT) of one instance: all cubes of width W in
  one persistent launch of the counting kernel, result x 2^free summed per cube in
  checked uint64; raises ``OverflowError`` instead of returning a wrapped count.
* :func:`count_search` — the same counts from the device cube search in exhaustive mode
  (no host cube enumeration), for instances of at most 128 variables and 3-literal
  clauses, several at once.

Counterpart of ``bool(CotengraContraction(FinRel, wiring).contract(diagram))``
(``examples/test_sat_solver.py:5-9``) and of the #SAT prototype's counting einsum
(``notebooks/prototypes/2025-02-13-#sat-tensor-contraction.ipynb`` cell 12), which has no
package function in the reference.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import program as PG
from .executor import BatchCounter, DeviceProgram, gang_stride, run_gang, run_job
from .sat import CNFInstance


@dataclass
class SingleResult:
    value: int
    plan_s: float
    gpu_s: float
    cubes: int
    cubes_run: int
    stats: dict = field(default_factory=dict)


def decide(cnf: CNFInstance, *, width: int = 26, mode: str = "GT0", plan_trials: int = 2,
           program: PG.Program | None = None, streams: int = 4, small_bits: int = 12, gang: int | None = None,
           first_gang: int = 16, lookahead: bool = False) -> SingleResult:
    """See the module docstring.  With ``gang`` > 0 the cubes (or slices) run in gangs
    (``ts_program_run_gang``) of ``first_gang`` jobs, doubling up to ``gang`` (None: 32
    when every job has the same steps -- the slices of one sliced tree, whose shared
    small-step runs cost the same for 1 or 32 jobs -- else 8, cubes of different shapes
    where launching past the first satisfiable one costs most; measured on the n=40
    slices and the n=120 cubes); with
    ``lookahead`` each gang is enqueued before the previous one's flags are read (the
    device never waits for the host, at the price of one more gang past the deciding
    one), else the flags are read between gangs.  ``gang=0`` runs the cubes one by one on
    ``streams`` streams.  ``stats["cubes_launched"]`` counts the cubes that ran (the rest
    of the deciding gang, and the gang after it with ``lookahead``); ``cubes_run`` those up
    to the deciding one."""
    import torch

    t0 = time.perf_counter()
    prog = program if program is not None else PG.sat_batch([cnf], width=width, plan_trials=plan_trials)
    plan_s = time.perf_counter() - t0
    if len(prog.jobs) == 0:
        return SingleResult(0, plan_s, 0.0, 0, 0, {"host_unsat": True})
    if gang is None:
        jobs = prog.jobs
        same = bool(np.all(jobs["num_steps"] == jobs["num_steps"][0]) and
                    np.all(jobs["arena_elems"] == jobs["arena_elems"][0]))
        gang = 32 if same else 8
    if streams < 1 or gang < 0 or first_gang < 1:
        raise ValueError("streams and first_gang must be >= 1, gang >= 0")
    dp = DeviceProgram(prog)
    stats = {"max_rank": int(prog.stats[0, 4]), "branch_nodes": int(prog.stats[0, 1])}
    if gang == 0:
        verdict, ran, launched, t = _decide_streams(dp, mode, streams, small_bits)
        stats.update(cubes_launched=launched, streams=min(streams, len(prog.jobs)))
    else:
        verdict, ran, launched, t = _decide_gangs(dp, mode, gang, first_gang, small_bits, lookahead)
        stats.update(cubes_launched=launched, gang=gang)
    return SingleResult(verdict, plan_s, t, len(prog.jobs), ran, stats)


def _decide_gangs(dp: DeviceProgram, mode: str, gang: int, first_gang: int, small_bits: int, lookahead: bool):
    import collections

    import torch

    nj = len(dp.jobs_host)
    size = min(first_gang, gang)
    arena = torch.empty(min(gang, nj) * gang_stride(dp), dtype=torch.uint8, device=dp.device)
    flags_dev = torch.zeros(nj, dtype=torch.uint8, device=dp.device)
    flags = torch.zeros(nj, dtype=torch.uint8, pin_memory=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    verdict, ran, begin = 0, 0, 0
    pending: collections.deque = collections.deque()
    while begin < nj or pending:
        if begin < nj and len(pending) < (2 if lookahead else 1):
            n = min(size, nj - begin)
            run_gang(dp, begin, n, mode, arena=arena, results=flags_dev[begin : begin + n], small_bits=small_bits)
            flags[begin : begin + n].copy_(flags_dev[begin : begin + n], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            pending.append((begin, n, ev))
            begin += n
            size = min(gang, 2 * size)
            continue
        b, n, ev = pending.popleft()
        ev.synchronize()
        hits = np.flatnonzero(flags[b : b + n].numpy())
        if hits.size:
            verdict, ran = 1, b + int(hits[0]) + 1
            break
        ran = b + n
    torch.cuda.synchronize()
    return verdict, ran, begin, time.perf_counter() - t0


def _decide_streams(dp: DeviceProgram, mode: str, streams: int, small_bits: int):
    import collections

    import torch

    from .executor import run_job_fused

    nj = len(dp.jobs_host)
    size = max(16, int(dp.jobs_host["arena_elems"].max()))
    lanes = [(torch.cuda.Stream(), torch.empty(size, dtype=torch.uint8, device=dp.device))
             for _ in range(min(streams, nj))]
    results = torch.zeros(nj, dtype=torch.uint8, device=dp.device)
    main = torch.cuda.current_stream()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    verdict, ran, launched = 0, 0, 0
    inflight: collections.deque = collections.deque()
    while True:
        while launched < nj and len(inflight) < len(lanes):
            st, arena = lanes[launched % len(lanes)]
            st.wait_stream(main)
            with torch.cuda.stream(st):
                res = run_job_fused(dp, launched, mode, arena=arena, small_bits=small_bits)
                results[launched : launched + 1].copy_(res.amax().reshape(1))
                ev = torch.cuda.Event()
                ev.record(st)
            inflight.append((launched, ev))
            launched += 1
        if not inflight:
            break
        j, ev = inflight.popleft()
        ev.synchronize()
        ran = j + 1
        if int(results[j].item()):
            verdict = 1
            break
    torch.cuda.synchronize()
    return verdict, ran, launched, time.perf_counter() - t0


def decide_budget(cnf: CNFInstance, *, width: int = 22, budget_s: float = 20.0, mode: str = "GT0",
                  plan_trials: int = 8, priority: np.ndarray | None = None, prior_branch: bool = False,
                  gate_margin: float = 0.5, first_chunk: int = 16, node_budget: int = 100_000,
                  threads: int = 1) -> SingleResult:
    """Budgeted decision (see the module docstring).  ``priority`` (learned scores) makes
    the search use the learned proposals; ``plan_s`` includes every host planning call."""
    import torch

    from .executor import BatchExecutor

    t_start = time.perf_counter()
    plan_s = gpu_s = 0.0
    begin, chunk, max_nodes = 0, first_chunk, node_budget
    cubes = 0
    stats = {"learned_trees": 0, "fallback_trees": 0, "branch_nodes": 0, "calls": 0}
    rate = None  # branch nodes per second of the host search, measured on the first call
    while True:
        left = budget_s - (time.perf_counter() - t_start)
        if left <= 0:
            return SingleResult(-1, plan_s, gpu_s, cubes, cubes, stats)
        if rate:
            # keep one call inside the remaining budget (the search re-walks earlier nodes)
            max_nodes = max(1000, min(max_nodes, int(rate * left)))
        t0 = time.perf_counter()
        prog = PG.sat_batch([cnf], width=width, cube_begin=begin, max_cubes=chunk, max_nodes=max_nodes,
                            stop_on_sat=True, plan_trials=plan_trials, priority=priority, prior_branch=prior_branch,
                            gate_margin=gate_margin, threads=threads)
        dt = time.perf_counter() - t0
        plan_s += dt
        st = prog.stats[0]
        if st[1] > 1000 and dt > 0:
            rate = st[1] / dt
        stats["calls"] += 1
        stats["branch_nodes"] = max(stats["branch_nodes"], int(st[1]))
        stats["learned_trees"] += int(st[8])
        stats["fallback_trees"] += int(st[9])
        cubes += len(prog.jobs)
        if st[5] == 1:
            return SingleResult(1, plan_s, gpu_s, cubes, cubes, dict(stats, decided="host: satisfied residual"))
        if st[5] == 0:
            return SingleResult(0, plan_s, gpu_s, cubes, cubes, dict(stats, decided="host: every branch refuted"))
        if len(prog.jobs):
            t0 = time.perf_counter()
            v = int(BatchExecutor(prog).run(mode)[0])
            torch.cuda.synchronize()
            gpu_s += time.perf_counter() - t0
            if v:
                return SingleResult(1, plan_s, gpu_s, cubes, cubes, dict(stats, decided="gpu: satisfiable cube"))
        if st[3] == 0:
            return SingleResult(0, plan_s, gpu_s, cubes, cubes, dict(stats, decided="all cubes unsatisfiable"))
        if st[3] == 2:
            max_nodes *= 2  # the search ran out of branch nodes before filling the chunk
        begin += len(prog.jobs)
        chunk *= 2


def decide_learned(cnf: CNFInstance, priority: np.ndarray, *, width: int = 22, budget_s: float = 20.0,
                   slice_s: float = 0.25, mode: str = "GT0", plan_trials: int = 8,
                   gate_margin: float = 0.5) -> SingleResult:
    """
    The learned planner with the plain planner as fallback, under one budget: time
    slices of ``decide_budget`` alternate between the learned search (learned branching
    and tree proposals) and the plain search, each slice twice as long as the previous
    pair's.  Every slice restarts its search, so the answer is whichever search decides
    first; both are exact.
    """
    t_start = time.perf_counter()
    plan_s = gpu_s = 0.0
    cubes = 0
    stats = {"learned_trees": 0, "fallback_trees": 0, "slices": 0}
    s = slice_s
    while True:
        for learned in (True, False):
            left = budget_s - (time.perf_counter() - t_start)
            if left <= 0:
                return SingleResult(-1, plan_s, gpu_s, cubes, cubes, stats)
            kw = dict(priority=priority, prior_branch=True) if learned else {}
            r = decide_budget(cnf, width=width, budget_s=min(s, left), mode=mode, plan_trials=plan_trials,
                              gate_margin=gate_margin, **kw)
            plan_s += r.plan_s
            gpu_s += r.gpu_s
            cubes += r.cubes
            stats["slices"] += 1
            stats["learned_trees"] += r.stats["learned_trees"]
            stats["fallback_trees"] += r.stats["fallback_trees"]
            if r.value >= 0:
                return SingleResult(r.value, plan_s, gpu_s, cubes, cubes,
                                    dict(stats, decided=r.stats.get("decided"), by="learned" if learned else "plain"))
        s *= 2


def count_search(cnfs, *, width: int = 6, max_split: int = 10, budget: int | None = 4096) -> SingleResult:
    """Exact model counts of CNF instances (at most 128 variables, 3 literals per clause)
    by the device cube search in exhaustive mode: every instance split into 2^max_split
    start jobs, each leaf (residual of at most ``width`` variables) adding its relation's
    entry count times 2 to the number of variables no residual clause holds (the
    counting einsum of the #SAT prototype, with the search's conditioning in place of
    the host's cube enumeration).  ``budget``: nodes a job searches before it hands its
    remaining subtrees over (None: no hand-over; the default balanced the n=100 search
    best).  ``value``: the list of counts; raises OverflowError for a count beyond 64
    bits.  Pays for batches of small instances; at n = 100, m = 300 the width-6 search
    tree is far larger than the width-18 cube list of :func:`count` (4.6 s against 1.5 s
    per instance, planning included)."""
    import torch

    from .search import SearchSolver

    cnfs = list(cnfs) if not isinstance(cnfs, CNFInstance) else [cnfs]
    t0 = time.perf_counter()
    solver = SearchSolver(width=width, split_target=1e-9, max_split=max_split, budget=budget, count=True,
                          max_nodes=1 << 40)
    solver.prepare(cnfs)
    plan_s = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.launch()
    counts = solver.counts()
    gpu_s = time.perf_counter() - t0
    jobs = len(solver.host_jobs)
    return SingleResult(counts, plan_s, gpu_s, jobs, jobs, {"rounds": solver.rounds, "width": width,
                                                            "max_split": max_split})


def count(cnf: CNFInstance, *, width: int = 18, plan_trials: int = 1,
          program: PG.Program | None = None) -> SingleResult:
    import torch

    t0 = time.perf_counter()
    prog = program if program is not None else PG.sat_batch([cnf], width=width, plan_trials=plan_trials)
    plan_s = time.perf_counter() - t0
    if len(prog.jobs) == 0:
        return SingleResult(0, plan_s, 0.0, 0, 0, {"host_unsat": True})
    bc = BatchCounter(prog)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    (c,) = bc.run()
    return SingleResult(c, plan_s, time.perf_counter() - t0, len(prog.jobs), len(prog.jobs),
                        {"max_rank": int(prog.stats[0, 4]), "branch_nodes": int(prog.stats[0, 1])})
