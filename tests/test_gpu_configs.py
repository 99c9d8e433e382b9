"""
GPU parity on the measurement configurations at their own sizes:

* the plain n=40, m=170 instances decided through the reference's own call shape
  (``examples/test_sat_solver.py:5-9``: a planned contraction of the flattened SAT
  diagram, here ``PlannedContraction(DeviceRel, wiring, optimize="greedy")`` sliced to
  2^26-entry intermediates), against DPLL and the oracle's verdict on a cube program;
* every intermediate of every width-26 cube of the n=120, m=510 instance against the
  oracle's step loop, by per-cube digest chains committed in ``tests/golden``;
* the exact n=100, m=300 model counts against an independent model counter's
  (``oracle/sharp.c``) committed counts;
* learned-planner decisions (n = 150..300) of both the plain and the learned search
  against DPLL.
"""

import json
from pathlib import Path

import numpy as np
import pytest
import xxhash

from oracle import dpll
from oracle import executor as OX
from tensorsat_amd import program as PG
from tensorsat_amd.workloads import count_3sat, learned_ksat, plain_cpu_3sat, single_3sat

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden"


@pytest.mark.parametrize("seed", range(10))
def test_plain_n40_greedy_drop_in(seed):
    from tensorsat_amd.contractions import PlannedContraction
    from tensorsat_amd.rel import DeviceRel

    cnf = plain_cpu_3sat(seed)
    d = (cnf.inputs(None) >> cnf.diagram()).flatten()
    c = PlannedContraction(DeviceRel, d.wiring, optimize="greedy", slicing_target=26)
    print(f"seed {seed}: {len(c.tree.sliced)} sliced wires, max 2^{c.tree.cost.log2_max_size:.0f}", flush=True)
    got = bool(c.contract(d))
    want = dpll.solve(cnf.num_vars, cnf.clauses)
    assert got == want
    # the oracle's verdict on a cube decomposition of the same instance
    prog = PG.sat_batch([cnf], width=10, plan_trials=1)
    assert bool(OX.decide(prog)[0]) == want


@pytest.mark.parametrize("seed", range(3))
def test_plain_n40_greedy_compiled_slices(seed):
    """The same greedy trees (sliced to 2^26) compiled one job per slice and decided by the
    step executors (config 1 as the bench measures it): the verdict equals DPLL's, and
    every intermediate of the first slice equals the oracle's step loop on the same
    records, byte for byte."""
    from tensorsat_amd.contractions import PlannedContraction
    from tensorsat_amd.executor import DeviceProgram, run_job
    from tensorsat_amd.rel import DeviceRel
    from tensorsat_amd.single import decide

    cnf = plain_cpu_3sat(seed)
    d = (cnf.inputs(None) >> cnf.diagram()).flatten()
    c = PlannedContraction(DeviceRel, d.wiring, optimize="greedy", slicing_target=26)
    prog = c.contract_program(d)
    assert len(prog.jobs) == 1 << len(c.tree.sliced)
    r = decide(cnf, program=prog)
    want = dpll.solve(cnf.num_vars, cnf.clauses)
    print(f"seed {seed}: {len(prog.jobs)} slices, verdict {r.value} (dpll {want}), {r.cubes_run} run", flush=True)
    assert r.value == int(want)
    if seed == 0:
        res, inter = run_job(DeviceProgram(prog), 0, "GT0", keep_intermediates=True)
        ref, ref_inter = OX.run_job(prog, 0, "GT0", keep_intermediates=True)
        for s, (a, b) in enumerate(zip(inter, ref_inter)):
            np.testing.assert_array_equal(a.cpu().numpy(), b, err_msg=f"step {s}")
        np.testing.assert_array_equal(res.cpu().numpy(), ref)


def _chain(step_outputs, result) -> str:
    h = xxhash.xxh64()
    for t in step_outputs:
        h.update(xxhash.xxh64(t.cpu().numpy().tobytes()).digest())
    h.update(xxhash.xxh64(result.cpu().numpy().tobytes()).digest())
    return h.hexdigest()


def test_n120_width26_every_intermediate_digest():
    """All cubes of the n=120 instance at width 26 (the 2^26-entry grid steps): every
    intermediate's xxh64, chained per cube, equals the oracle's (GT0 for every cube,
    MOD256 for the first ones)."""
    from tensorsat_amd.executor import DeviceProgram, run_job

    want = json.loads((GOLDEN / "config2_w26_digests.json").read_text())
    prog = PG.sat_batch([single_3sat(want["seed"])], width=want["width"], plan_trials=want["plan_trials"],
                        threads=want["threads"])
    h = xxhash.xxh64()
    for a in (prog.jobs, prog.steps, prog.leaves):
        h.update(a.tobytes())
    assert h.hexdigest() == want["program"], "the planner's program changed: regenerate the fixture"
    dp = DeviceProgram(prog)
    for mode in ("GT0", "MOD256"):
        for j, ref in enumerate(want[mode]):
            res, inter = run_job(dp, j, mode, keep_intermediates=True)
            got = _chain(inter, res)
            print(f"{mode} cube {j}: {got}", flush=True)
            assert got == ref, f"{mode} cube {j}"


def test_n100_counts_match_independent_counter():
    from tensorsat_amd.single import count

    want = json.loads((GOLDEN / "sharp_n100_counts.json").read_text())["counts"]
    for seed in range(10):
        r = count(count_3sat(seed), width=18)
        print(f"seed {seed}: {r.value} ({r.cubes} cubes, gpu {r.gpu_s:.3f}s)", flush=True)
        assert r.value == int(want[str(seed)])


def test_learned_workload_decisions_both_arms():
    """Four k=3 instances of the learned-planner workload (held-out seeds): the plain
    budgeted search and the learned one (network branching and tree proposals, plain
    fallback) both decide them, with the verdict DPLL gives where it finishes."""
    from tensorsat_amd.learn.propose import LearnedPlanner
    from tensorsat_amd.single import decide_budget, decide_learned

    picks = [i for i in range(40) if max(len(c) for c in learned_ksat(i, base=2_000_000).clauses) == 3][:4]
    cnfs = [learned_ksat(i, base=2_000_000) for i in picks]
    scores = LearnedPlanner().scores(cnfs)
    for i, c, s in zip(picks, cnfs, scores):
        want = dpll.solve(c.num_vars, c.clauses, budget=20_000_000)
        plain = decide_budget(c, width=22, budget_s=30.0)
        learned = decide_learned(c, s, width=22, budget_s=30.0)
        print(f"i={i} n={c.num_vars} m={len(c.clauses)}: dpll {want} plain {plain.value} ({plain.plan_s:.2f}s plan) "
              f"learned {learned.value} ({learned.plan_s:.2f}s plan, by {learned.stats.get('by')})", flush=True)
        assert plain.value in (0, 1) and learned.value in (0, 1)
        assert plain.value == learned.value
        if want is not None:
            assert plain.value == int(want)
