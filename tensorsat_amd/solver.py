"""
Batched SAT decisions: conditioned slicing on the host, contraction on the device.

:class:`BatchSolver` decides many CNF instances at once.  The host front-end
(``program.sat_batch``) slices every instance into cubes of width <= W and compiles the
first ``first_pass`` cubes of each into contraction jobs; the device runs them in
persistent launches with early exit.  Instances still undecided whose cube list was cut
off get a second pass over the rest of their cubes, whose jobs are appended to the first
pass's program (one launch decides the batch).  Both passes are prepared once
(:meth:`prepare`), so :meth:`run` only launches kernels on device-resident jobs.

The verdict of an instance is the OR over all its cubes; every pass keeps that OR, so
the answer does not depend on the pass split.  This is the batched counterpart of the
reference's per-instance ``bool(CotengraContraction(...).contract(diagram))``
(``examples/test_sat_solver.py:5-9``).
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import program as PG
from .executor import BatchExecutor
from .sat import CNFInstance


@dataclass
class SolveStats:
    plan_s: float = 0.0
    pass1_jobs: int = 0
    pass2_jobs: int = 0
    pass2_instances: int = 0
    host_unsat: int = 0
    kernel_ms: list = field(default_factory=list)


class BatchSolver:
    def __init__(self, *, width: int = 10, first_pass: int = 4, mode: str = "GT0", threads: int = 0,
                 plan_trials: int = 1, num_workgroups: int | None = None, track_jobs: bool = True,
                 lds: bool | None = None, round_bounds=None, merge_passes: bool = True) -> None:
        """``merge_passes``: run the second pass's jobs in the first pass's launch (one
        program; default) rather than in a launch of their own after it."""
        self.width = width
        self.merge_passes = merge_passes
        self.first_pass = first_pass
        self.mode = mode
        self.threads = threads
        self.plan_trials = plan_trials
        self.num_workgroups = num_workgroups
        self.track_jobs = track_jobs
        self.lds = lds
        self.round_bounds = round_bounds
        self.stats = SolveStats()
        self.progs: list[PG.Program] = []
        self.execs: list[BatchExecutor] = []
        self.subsets: list[np.ndarray] = []

    def prepare(self, cnfs: list[CNFInstance]) -> None:
        """Plan and compile both passes (runs the first pass once to select the second)."""
        self.n = len(cnfs)
        t0 = time.perf_counter()
        p1 = PG.sat_batch(cnfs, width=self.width, max_cubes=self.first_pass, plan_trials=self.plan_trials,
                          threads=self.threads)
        self.stats.plan_s = time.perf_counter() - t0
        self.stats.pass1_jobs = len(p1.jobs)
        self.stats.host_unsat = int((p1.stats[:, 5] == 0).sum())
        rb = {} if self.round_bounds is None else {"round_bounds": tuple(self.round_bounds)}
        ex1 = BatchExecutor(p1, num_workgroups=self.num_workgroups, track_jobs=self.track_jobs, lds=self.lds, **rb)
        v1 = ex1.run(self.mode)
        self.progs = [p1]
        self.execs = [ex1]
        self.subsets = [np.arange(self.n)]
        need = np.flatnonzero((v1 == 0) & (p1.stats[:, 3] == 1))
        if need.size:
            t0 = time.perf_counter()
            p2 = PG.sat_batch([cnfs[i] for i in need], width=self.width, cube_begin=self.first_pass,
                              plan_trials=self.plan_trials, threads=self.threads)
            self.stats.plan_s += time.perf_counter() - t0
            self.stats.pass2_jobs = len(p2.jobs)
            self.stats.pass2_instances = int(need.size)
            if len(p2.jobs) and self.merge_passes:
                # One program, one launch: the second pass's cubes queue behind the first
                # pass's (dispatch order is by cube index) and share its early exit.
                merged = PG.concat([p1, p2], [np.arange(self.n), need], self.n)
                self.progs = [merged]
                self.execs = [BatchExecutor(merged, num_workgroups=self.num_workgroups,
                                            track_jobs=self.track_jobs, lds=self.lds, **rb)]
            elif len(p2.jobs):
                self.progs.append(p2)
                self.execs.append(BatchExecutor(p2, num_workgroups=self.num_workgroups,
                                                track_jobs=self.track_jobs, lds=self.lds, **rb))
                self.subsets.append(need)

    def launch(self, events=None) -> None:
        """Enqueue every pass on the current stream.  ``events`` (start, end): with one
        launch (passes merged) they bracket its persistent kernel alone, after the reset;
        with separate passes they bracket all of them, resets included."""
        if len(self.execs) == 1:
            self.execs[0].launch(self.mode, events=events)
            return
        if events is not None:
            events[0].record()
        for ex in self.execs:
            ex.launch(self.mode)
        if events is not None:
            events[1].record()

    def capture(self, steps: int = 1) -> None:
        """Record every pass (resets, compactions, persistent launches) into one HIP graph,
        ``steps`` complete decisions of the batch back to back; :meth:`replay` then runs
        them with a single graph launch (several decisions per graph amortise the host
        cost of a graph launch, which exceeds one decision's device time).  The graph
        holds kernel nodes only (the resets are fill kernels, not memsets)."""
        import torch

        if steps < 1:
            raise ValueError("steps must be >= 1")
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.launch()  # warm the allocator and the code objects outside capture
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            for _ in range(steps):
                self.launch()
        torch.cuda.synchronize()
        self.graph_steps = steps

    def replay(self, events=None) -> None:
        """One graph launch of the captured decisions (see :meth:`capture`)."""
        if events is not None:
            events[0].record()
        self.graph.replay()
        if events is not None:
            events[1].record()

    def verdicts(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.uint8)
        for ex, sub in zip(self.execs, self.subsets):
            v = ex.verdicts.cpu().numpy()[: len(sub)].astype(np.uint8)
            out[sub] |= v
        return out

    def run(self) -> np.ndarray:
        import torch

        self.launch()
        torch.cuda.synchronize()
        return self.verdicts()

    def executed_jobs(self) -> list[np.ndarray]:
        """Per pass: boolean mask of jobs that ran in the last launch (needs track_jobs)."""
        return [ex.job_flags() == 1 for ex in self.execs]


def autotune_width(cnfs: list[CNFInstance], *, widths=(6, 8, 10, 12), sample: int = 512, first_pass: int = 4,
                   mode: str = "GT0", reps: int = 3, threads: int = 0) -> tuple[int, dict]:
    """
    Pick the cube width W for a batch: on the first ``sample`` instances, each candidate
    is planned (host time) and decided ``reps`` times on the GPU (device time); the width
    with the smallest plan + decide time wins.  Returns (W, {W: measurements}).
    Smaller W moves work from the contraction to the host search (more branches refuted
    by propagation, smaller residuals); larger W means fewer, wider cubes.
    """
    import torch

    sub = cnfs[:sample]
    table = {}
    for w in widths:
        s = BatchSolver(width=w, first_pass=first_pass, mode=mode, threads=threads, track_jobs=False)
        s.prepare(sub)
        s.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            s.launch()
        torch.cuda.synchronize()
        gpu_s = (time.perf_counter() - t0) / reps
        table[w] = {"plan_s": s.stats.plan_s, "gpu_ms": gpu_s * 1000.0, "cost_s": s.stats.plan_s + gpu_s,
                    "jobs": s.stats.pass1_jobs + s.stats.pass2_jobs, "host_unsat": s.stats.host_unsat}
    best = min(table, key=lambda w: table[w]["cost_s"])
    return best, table
