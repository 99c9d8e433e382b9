"""Planner (libtsplan.so): path validity, cost model closed forms, slicing, cube front-end."""

import itertools
import math

import numpy as np
import pytest

from oracle import brute
from oracle import executor as OX
from tensorsat_amd import plan as P
from tensorsat_amd import program as PG
from tensorsat_amd.network import Wiring, bintree_network, clause_network
from tensorsat_amd.sat import CNFInstance


def _valid_ssa(plan, S):
    alive = set(range(S))
    for k, (a, b) in enumerate(plan.path.tolist()):
        assert a in alive and b in alive and a != b
        alive -= {a, b}
        alive.add(S + k)
    assert len(alive) == 1


@pytest.mark.parametrize("planner", ["greedy", "hyper", "elim"])
def test_paths_are_complete_trees(planner):
    cnf = CNFInstance.random(3, 20, 85, rng=3)
    for d in (bintree_network(cnf), clause_network(cnf)):
        if planner == "greedy":
            pl = P.greedy(d.wiring)
        elif planner == "hyper":
            pl = P.hyper(d.wiring, trials=8, threads=2)
        else:
            pl = PG.elim(d.wiring, trials=4)
        _valid_ssa(pl, d.wiring.num_slots)


def test_cost_model_closed_form():
    # chain A(ab) B(bc) C(cd), out = (a, d): merge (A,B) then (AB, C)
    w = Wiring((2, 2, 2, 2), [(0, 1), (1, 2), (2, 3)], [0, 3])
    plan = P.Plan(3, np.array([[0, 1], [2, 3]], dtype=np.int32))
    c = P.tree_cost(w, plan)
    # step 1 union {a,b,c}: 8, out {a,c}: 4; step 2 union {a,c,d}: 8, out {a,d}: 4
    assert math.isclose(c.log2_flops, math.log2(16))
    assert math.isclose(c.log2_write, math.log2(8))
    assert c.log2_max_size == 2
    # peak: leaves 4+4+4=12, +4 (AB) = 16, then 16-8 = 8 live, +4 = 12
    assert math.isclose(c.log2_peak, math.log2(16))
    # extent-3 wires
    w3 = Wiring((3, 3, 3), [(0, 1), (1, 2)], [])
    c3 = P.tree_cost(w3, P.Plan(2, np.array([[0, 1]], dtype=np.int32)))
    assert math.isclose(c3.log2_flops, math.log2(27)) and c3.log2_max_size == 0


def test_linear_path_roundtrip():
    cnf = CNFInstance.random(3, 12, 40, rng=0)
    d = clause_network(cnf)
    pl = P.greedy(d.wiring)
    lp = pl.linear_path()
    back = P.Plan.from_linear_path(d.wiring.num_slots, lp)
    np.testing.assert_array_equal(back.path, pl.path)


def test_slicing_reaches_target_and_stays_exact():
    cnf = CNFInstance.random(3, 18, 60, rng=7)
    d = clause_network(cnf)
    pl = PG.elim(d.wiring, trials=2)
    target = pl.cost.log2_max_size - 3
    sl = P.slice_plan(d.wiring, pl, target, reconfigure_trials=8)
    assert sl.cost.log2_max_size <= target + 1e-9
    assert len(sl.sliced) >= 1
    # summing the sliced contractions equals the brute-force count
    total = 0
    base = [np.asarray(b) for b in d.blocks]
    for vals in itertools.product(range(2), repeat=len(sl.sliced)):
        fixed = dict(zip(sl.sliced, vals))
        sw, leaves = [], []
        for ws, t in zip(d.wiring.slot_wires_list, base):
            idx = tuple(fixed.get(w, slice(None)) for w in ws)
            sw.append(tuple(w for w in ws if w not in fixed))
            leaves.append(np.ascontiguousarray(t[idx]))
        keep = [w for w in range(d.wiring.num_wires) if w not in fixed]
        remap = {w: i for i, w in enumerate(keep)}
        wr = Wiring((2,) * len(keep), [tuple(remap[w] for w in ws) for ws in sw], [])
        prog = PG.compile_network(wr, leaves, PG.elim(wr, trials=1))
        res, _ = OX.run_job(prog, 0, "COUNT")
        total += int(res.reshape(-1)[0])
    assert total == brute.count(18, cnf.clauses)


def test_invalid_inputs_raise():
    w = Wiring((2, 2), [(0, 1), (1,)], [])
    with pytest.raises(ValueError):
        P.tree_cost(w, P.Plan(2, np.array([[0, 0]], dtype=np.int32)))


@pytest.mark.parametrize("width", [4, 8, 12])
def test_cube_front_end_exact(width):
    cnfs = [CNFInstance.random(3, n, int(round(r * n)), rng=50 + n) for n in (14, 18, 22) for r in (3.0, 4.26)]
    prog = PG.sat_batch(cnfs, width=width)
    PG.validate(prog)
    assert prog.stats[:, 4].max() <= max(width + 4, 8)  # residual width estimate bounds the job rank
    counts = OX.count(prog)
    for c, got in zip(cnfs, counts):
        assert got == brute.count(c.num_vars, c.clauses)


def test_cube_passes_partition_the_cubes():
    cnfs = [CNFInstance.random(3, 22, 80, rng=s) for s in range(4)]
    full = PG.sat_batch(cnfs, width=6)
    p1 = PG.sat_batch(cnfs, width=6, max_cubes=2)
    p2 = PG.sat_batch(cnfs, width=6, cube_begin=2)
    c_full = OX.count(full)
    c_split = [a + b for a, b in zip(OX.count(p1), OX.count(p2))]
    assert c_full == c_split
    assert len(p1.jobs) + len(p2.jobs) == len(full.jobs)


@pytest.mark.parametrize("width", [6, 10])
def test_parallel_enumeration_reproduces_sequential_program(width):
    """One instance, full enumeration: the search tree is split at a fixed depth and the
    subtrees are enumerated by 8 threads; concatenated in order they give exactly the
    sequential program (deterministic trees with one planning trial)."""
    from tensorsat_amd.workloads import count_3sat

    cnf = count_3sat(1)
    seq = PG.sat_batch([cnf], width=width, plan_trials=1, threads=1)
    par = PG.sat_batch([cnf], width=width, plan_trials=1, threads=8)
    assert len(seq.jobs) == len(par.jobs) > 64
    for f in ("num_steps", "num_leaf_bytes", "result_rank", "free_vars", "arena_elems", "cube"):
        np.testing.assert_array_equal(seq.jobs[f], par.jobs[f])
    np.testing.assert_array_equal(seq.steps.view(np.uint8), par.steps.view(np.uint8))
    np.testing.assert_array_equal(seq.leaves, par.leaves)
    assert seq.stats[0, 1] == par.stats[0, 1]  # branch nodes


def test_concat_programs_keeps_every_verdict():
    """Two programs over overlapping instance sets, concatenated with renamed instances:
    valid (aligned leaf blocks, in-range steps) and decided like the OR of the parts."""
    from oracle import executor as OX
    from tensorsat_amd.workloads import batched_3sat

    cnfs = batched_3sat(40, start=300)
    p1 = PG.sat_batch(cnfs, width=6, max_cubes=2)
    sub = np.arange(0, 40, 3)
    p2 = PG.sat_batch([cnfs[i] for i in sub], width=6, cube_begin=2)
    p1.leaves = np.concatenate([p1.leaves, np.ones(5, np.uint8)])  # an unaligned tail to pad
    m = PG.concat([p1, p2], [np.arange(40), sub], 40)
    PG.validate(m)
    assert len(m.jobs) == len(p1.jobs) + len(p2.jobs)
    want = OX.decide(p1, "GT0")
    want[sub] |= OX.decide(p2, "GT0")
    np.testing.assert_array_equal(OX.decide(m, "GT0"), want)
    np.testing.assert_array_equal(m.stats, p1.stats)
    with pytest.raises(ValueError):
        PG.concat([p2], [sub + 40], 40)


def test_planner_scratch_reuse_leaves_no_state():
    """The planner and the step compiler keep per-thread scratch from cube to cube and
    call to call: a program compiled after larger programs on the same thread, or inside
    a batch whose other instances run on the same worker threads, is byte-identical to
    the same program compiled first and alone."""
    from tensorsat_amd.workloads import single_3sat

    a = CNFInstance.random(3, 60, 200, rng=3)
    b = single_3sat(0)
    first = PG.sat_batch([a], width=8, plan_trials=1, threads=1)
    big = PG.sat_batch([b], width=20, plan_trials=1, threads=1)  # wider residuals, longer jobs
    assert int(big.jobs["num_steps"].max()) > int(first.jobs["num_steps"].max())
    again = PG.sat_batch([a], width=8, plan_trials=1, threads=1)
    assert len(first.jobs) == len(again.jobs) > 16
    np.testing.assert_array_equal(first.jobs, again.jobs)
    np.testing.assert_array_equal(first.steps.view(np.uint8), again.steps.view(np.uint8))
    np.testing.assert_array_equal(first.leaves, again.leaves)
    mixed = PG.sat_batch([b, a, b, a], width=8, plan_trials=1, threads=2)
    for inst in (1, 3):
        js = np.flatnonzero(mixed.jobs["instance"] == inst)
        assert len(js) == len(first.jobs)
        for k, j in enumerate(js):
            np.testing.assert_array_equal(mixed.job_steps(j).view(np.uint8), first.job_steps(k).view(np.uint8))
            np.testing.assert_array_equal(mixed.job_leaves(j), first.job_leaves(k))


def test_sliced_compilation_equals_per_slice_compile_network():
    """PlannedContraction.contract_program compiles one job per slice with the network,
    path and untouched leaves shared (program.compile_sliced): every job equals
    compile_network of that slice's sub-wiring and cut leaves, in product order."""
    from tensorsat_amd.contractions import PlannedContraction
    from tensorsat_amd.rel import DeviceRel

    cnf = CNFInstance.random(3, 20, 80, rng=7)
    d = (cnf.inputs(None) >> cnf.diagram()).flatten()
    c = PlannedContraction(DeviceRel, d.wiring, optimize="greedy", slicing_target=9)
    sliced = c.tree.sliced
    assert 2 <= len(sliced) <= 8
    prog = c.contract_program(d)
    assert len(prog.jobs) == 1 << len(sliced)
    w = d.wiring
    leaves = [np.asarray(getattr(b, "tensor", b), dtype=np.uint8) for b in d.blocks]
    sub = Wiring(w.wire_dims, [tuple(x for x in ws if x not in set(sliced)) for ws in w.slot_wires_list], ())
    for k, vals in enumerate(itertools.product(range(2), repeat=len(sliced))):
        fixed = dict(zip(sliced, vals))
        cut = [np.ascontiguousarray(t[tuple(fixed.get(x, slice(None)) for x in ws)])
               for t, ws in zip(leaves, w.slot_wires_list)]
        one = PG.compile_network(sub, cut, c.tree)
        np.testing.assert_array_equal(prog.job_steps(k).view(np.uint8), one.steps.view(np.uint8))
        np.testing.assert_array_equal(prog.job_leaves(k), one.leaves)
        for f in ("num_steps", "result_off", "result_rank", "arena_elems", "max_rank"):
            assert prog.jobs[f][k] == one.jobs[f][0]
