"""
GPU parity of the device cube search (``ts_search_run``) against the oracle's
restatement (``oracle/search.c``), the independent DPLL and brute force: verdicts of
the batched decision workload, exact per-instance statistics (branch nodes, leaves,
conflicts and the digest of every contracted leaf relation) with and without splitting,
edge cases (empty, unit, duplicate-literal and tautological clauses, the size limits),
and the node budget.
"""

import numpy as np
import pytest

from oracle import brute, dpll
from oracle import search as OS
from tensorsat_amd.sat import CNFInstance
from tensorsat_amd.workloads import batched_3sat

pytestmark = pytest.mark.gpu

FIELDS = ("nodes", "leaves", "conflicts", "digest")


@pytest.fixture(scope="module")
def batch():
    return batched_3sat(4096)


def _solve(cnfs, **kw):
    from tensorsat_amd.search import SearchSolver

    s = SearchSolver(**kw)
    v = s.run(cnfs)
    return s, v


def test_batched_workload_verdicts(batch):
    """All 4096 instances of the batched workload, split and launched hardest first with
    early exit: verdicts equal the oracle's and DPLL's."""
    s, v = _solve(batch, width=6, split_target=128.0)
    print(f"jobs {len(s.host_jobs)}, max split {int(s.host_jobs['split'].max())}, sat {int(v.sum())}", flush=True)
    ref = OS.search(batch, width=6, threads=8)["verdict"]
    np.testing.assert_array_equal(v, ref)
    want = np.array([dpll.solve(c.num_vars, c.clauses) for c in batch], dtype=np.int64)
    np.testing.assert_array_equal(v, want)


@pytest.mark.parametrize("ahead", [0, 1, 2])
def test_streamed_batches_match_one_at_a_time(batch, ahead):
    """decide_stream over batches of different sizes (the third needs a larger staging
    buffer than the first of its solver, the last ones are smaller again, and more
    batches than solvers in the ring), each searched while the next ones are staged on
    worker threads: every batch's verdicts equal an unstreamed solve and DPLL."""
    from tensorsat_amd.search import decide_stream

    parts = [batch[:300], batch[300:317], batch[317:3317], batch[3317:3400], batch[:1], batch[3400:4096]]
    got = list(decide_stream(parts, ahead=ahead, width=4, split_target=8.0))
    assert len(got) == len(parts)
    for k, (p, v) in enumerate(zip(parts, got)):
        _, want = _solve(p, width=4, split_target=8.0)
        print(f"batch {k}: {len(p)} instances, sat {int(v.sum())}", flush=True)
        np.testing.assert_array_equal(v, want)
        np.testing.assert_array_equal(v, np.array([dpll.solve(c.num_vars, c.clauses) for c in p], dtype=np.int64))
    with pytest.raises(ValueError):
        list(decide_stream(parts, width=4, budget=8))


@pytest.mark.parametrize("width", [0, 3, 6])
def test_unsplit_early_exit_statistics_exact(batch, width):
    """One job per instance: a sequential search, so every statistic equals the oracle's."""
    cnfs = batch[:1024]
    s, v = _solve(cnfs, width=width, split_target=None, stats=True)
    got = s.instance_stats()
    ref = OS.search(cnfs, width=width, threads=8)
    np.testing.assert_array_equal(v, ref["verdict"])
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
    np.testing.assert_array_equal(got["max_depth"], ref["max_depth"])


@pytest.mark.parametrize("target,budget,queue", [(None, None, False), (64.0, None, False), (4.0, None, False),
                                                (None, 4, False), (4.0, 2, False), (None, 1, False),
                                                (None, 4, True), (4.0, 2, True), (None, 1, True)])
def test_split_exhaustive_statistics_sum_to_one_search(batch, target, budget, queue):
    """Exhaustive searches (every branch visited): the jobs of an instance own disjoint
    subtrees -- static splits and the subtrees handed over when a job's budget runs out,
    to the next launch (rounds) or to the same launch's queue -- so their statistics add
    up to one unsplit search's, whatever the split."""
    cnfs = [c for c in batch[:600] if c.num_vars <= 50]
    s, v = _solve(cnfs, width=6, split_target=target, max_split=6, exhaustive=True, budget=budget, queue=queue)
    got = s.instance_stats()
    ref = OS.search(cnfs, width=6, exhaustive=True, threads=8)
    extra = f", queue records {s.queue_records()}" if s.queue else ""
    print(f"target {target} budget {budget} queue {queue}: rounds {s.rounds}, jobs per round {s.round_jobs[:8]}"
          f"{extra}", flush=True)
    np.testing.assert_array_equal(v, ref["verdict"])
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], ref[f], err_msg=f"{f} target {target} budget {budget}")
    if target == 4.0:
        assert int(s.host_jobs["split"].max()) >= 3
    if budget is not None and not queue:
        assert s.rounds > 3
    if queue:
        assert s.queue_records() > 2 * len(s.host_jobs)  # hand-overs ran in the one launch


@pytest.mark.parametrize("queue", [False, True])
def test_hand_over_list_overflow_keeps_every_subtree(batch, queue):
    """A continuation list (or queue) too small for the hand-overs: jobs that cannot hand
    over keep searching their subtrees themselves, so no branch is lost (exact
    exhaustive sums)."""
    cnfs = [c for c in batch[:300] if c.num_vars <= 45]
    cap = 16 if not queue else len(cnfs) + 16
    s, v = _solve(cnfs, width=6, split_target=None, exhaustive=True, budget=2, spill_capacity=cap, queue=queue)
    got = s.instance_stats()
    ref = OS.search(cnfs, width=6, exhaustive=True, threads=8)
    np.testing.assert_array_equal(v, ref["verdict"])
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
    if queue:
        assert s.queue_records() > cap  # more reserved than fit: searched in place
    else:
        assert s.rounds >= 2


@pytest.mark.parametrize("budget,queue", [(1, False), (8, False), (64, False), (1, True), (8, True), (64, True)])
def test_batched_workload_verdicts_with_hand_over(batch, budget, queue):
    s, v = _solve(batch, width=6, split_target=8.0, budget=budget, queue=queue)
    ref = OS.search(batch, width=6, threads=8)["verdict"]
    print(f"budget {budget}: rounds {s.rounds}, jobs per round {s.round_jobs[:10]}", flush=True)
    np.testing.assert_array_equal(v, ref)


def test_edge_cases_match_brute_force():
    cases = [
        CNFInstance(0, []),
        CNFInstance(3, []),
        CNFInstance(3, [()]),
        CNFInstance(3, [(1,), (-1,)]),
        CNFInstance(3, [(1,), (-1, 2), (-2, 3), (-3,)]),
        CNFInstance(4, [(1, 1, 2), (-1, -1), (2, -2, 3), (-2, -2, -2)]),
        CNFInstance(2, [(1, -1)]),
        CNFInstance(5, [(1, 2), (-1, 2), (1, -2), (-1, -2), (3, 4, 5)]),
    ]
    rng = np.random.default_rng(77)
    for _ in range(200):
        n = int(rng.integers(1, 15))
        cls = []
        for _ in range(int(rng.integers(0, 5 * n))):
            k = int(rng.choice([1, 2, 3, 3, 3]))
            cls.append(tuple(int(rng.integers(1, n + 1)) * (1 if rng.random() < 0.5 else -1) for _ in range(k)))
        cases.append(CNFInstance(n, cls))
    want = np.array([brute.is_sat(c.num_vars, c.clauses) for c in cases], dtype=np.int64)
    for width in (0, 6):
        for target in (None, 2.0):
            s, v = _solve(cases, width=width, split_target=target, max_split=4, stats=True)
            np.testing.assert_array_equal(v, want, err_msg=f"width {width} target {target}")
            if target is None:
                ref = OS.search(cases, width=width)
                got = s.instance_stats()
                for f in FIELDS:
                    np.testing.assert_array_equal(got[f], ref[f], err_msg=f)


def test_size_limits_128_vars_768_clauses():
    """The largest instances the kernel takes (all 12 clause slots of every lane, the
    upper half of the variable masks) against the oracle."""
    cnfs = [CNFInstance.random(3, 128, 768, rng=s) for s in range(4)]
    cnfs += [CNFInstance.random(3, 128, 450, rng=s) for s in range(4)]
    s, v = _solve(cnfs, width=6, split_target=None, stats=True, max_nodes=1 << 22)
    ref = OS.search(cnfs, width=6)
    np.testing.assert_array_equal(v, ref["verdict"])
    got = s.instance_stats()
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], ref[f], err_msg=f)


def test_node_budget_reports_undecided():
    """A job that runs out of its node budget leaves an UNSAT instance undecided (-1),
    never a wrong verdict."""
    cnfs = [c for c in batched_3sat(64) if c.num_vars >= 70][:6]
    ref = OS.search(cnfs, width=6)
    s, v = _solve(cnfs, width=6, split_target=None, max_nodes=3, stats=True)
    for i, c in enumerate(cnfs):
        if int(ref["nodes"][i]) > 3:
            assert v[i] in (-1, 1)
            if v[i] == 1:
                assert ref["verdict"][i] == 1
        else:
            assert v[i] == ref["verdict"][i]
    assert (v == -1).any()
