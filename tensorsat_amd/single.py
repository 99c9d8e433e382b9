"""
One large instance on the whole GPU.

* :func:`decide` — a SAT decision for one instance whose intermediates need the whole
  device (the n=120, m=510 workload: cubes of width up to 26, i.e. 2^26-entry tensors):
  cubes run one after the other in depth-first order, every step grid-wide, stopping at
  the first satisfiable cube.
* :func:`count` — the exact model count (#SAT) of one instance: all cubes of width W in
  one persistent launch of the counting kernel, result x 2^free summed per cube in
  checked uint64; raises ``OverflowError`` instead of returning a wrapped count.

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
from .executor import BatchCounter, DeviceProgram, run_job
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
           program: PG.Program | None = None) -> SingleResult:
    import torch

    t0 = time.perf_counter()
    prog = program if program is not None else PG.sat_batch([cnf], width=width, plan_trials=plan_trials)
    plan_s = time.perf_counter() - t0
    if len(prog.jobs) == 0:
        return SingleResult(0, plan_s, 0.0, 0, 0, {"host_unsat": True})
    dp = DeviceProgram(prog)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    verdict = 0
    ran = 0
    for j in range(len(prog.jobs)):
        res, _ = run_job(dp, j, mode)
        ran += 1
        if bool(res.any().item()):
            verdict = 1
            break
    torch.cuda.synchronize()
    return SingleResult(verdict, plan_s, time.perf_counter() - t0, len(prog.jobs), ran,
                        {"max_rank": int(prog.stats[0, 4]), "branch_nodes": int(prog.stats[0, 1])})


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
