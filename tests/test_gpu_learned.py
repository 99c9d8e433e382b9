"""
The trained planner network on the GPU: its proposals never change an answer, and what it
gains over the plain planner on the learned-planner workload (n = 150..300), measured as
branch nodes the cube search needs before it has a fixed number of contractible cubes.
"""

import numpy as np
import pytest

from oracle import dpll
from tensorsat_amd import program as PG
from tensorsat_amd.sat import CNFInstance
from tensorsat_amd.workloads import learned_ksat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def planner():
    from tensorsat_amd.learn.propose import LearnedPlanner

    return LearnedPlanner()


def test_scores_are_batch_independent(planner):
    cnfs = [learned_ksat(i, base=3_000_000) for i in range(4)]
    batch = planner.scores(cnfs)
    for c, s in zip(cnfs, batch):
        assert s.shape == (c.num_vars,) and np.isfinite(s).all()
        np.testing.assert_allclose(planner.scores([c])[0], s, rtol=1e-5, atol=1e-5)


def test_learned_proposals_keep_verdicts(planner):
    """Decisions with learned trees and learned branching equal the plain planner's and
    the independent DPLL's on instances small enough to finish."""
    from tensorsat_amd.single import decide_budget

    cnfs = [CNFInstance.random(3, n, int(round(r * n)), rng=700 + n) for n in (40, 50, 60) for r in (3.5, 4.5)]
    scores = planner.scores(cnfs)
    for c, s in zip(cnfs, scores):
        want = int(dpll.solve(c.num_vars, c.clauses))
        plain = decide_budget(c, width=12, budget_s=60.0)
        learned = decide_budget(c, width=12, budget_s=60.0, priority=s, prior_branch=True)
        print(f"n={c.num_vars} m={len(c.clauses)}: dpll {want} plain {plain.value} learned {learned.value} "
              f"(trees {learned.stats['learned_trees']} learned / {learned.stats['fallback_trees']} fallback)",
              flush=True)
        assert plain.value == want and learned.value == want
    # the exact cost model gate keeps part of the proposals
    assert learned.stats["learned_trees"] + learned.stats["fallback_trees"] > 0


def test_learned_branching_reaches_cubes_with_fewer_nodes(planner):
    """k=3 instances of the learned workload (n = 182..298): with the network's branching
    order the search fills 256 contractible cubes of width 22; the plain heuristic needs
    more branch nodes, or does not get there within 200k nodes."""
    picks = [i for i in range(30) if max(len(c) for c in learned_ksat(i, base=3_000_000).clauses) == 3][:6]
    cnfs = [learned_ksat(i, base=3_000_000) for i in picks]
    scores = planner.scores(cnfs)
    wins = 0
    for i, c, s in zip(picks, cnfs, scores):
        kw = dict(width=22, max_cubes=256, max_nodes=200_000, threads=1, plan_trials=8)
        plain = PG.sat_batch([c], **kw).stats[0]
        learned = PG.sat_batch([c], priority=s, prior_branch=True, gate_margin=0.5, **kw).stats[0]
        # nodes to fill the cube quota (the budget when the quota was not reached)
        cost_p = plain[1] if plain[0] >= 256 else 200_000 * 256 / max(plain[0], 1)
        cost_l = learned[1] if learned[0] >= 256 else 200_000 * 256 / max(learned[0], 1)
        print(f"i={i} n={c.num_vars}: plain {plain[0]} cubes / {plain[1]} nodes, learned {learned[0]} cubes / "
              f"{learned[1]} nodes", flush=True)
        wins += cost_l < cost_p
    assert wins >= len(picks) // 2
