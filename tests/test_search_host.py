"""
Host side of the device cube search: clause packing (normalisation as the oracle
restates it), job planning (every subtree owned exactly once) and the bounds checks that
run before any launch.  No GPU needed.
"""

import numpy as np
import pytest

from tensorsat_amd import search as S
from tensorsat_amd.sat import CNFInstance
from tensorsat_amd.workloads import batched_3sat


def _unpack(word):
    n = int(word) >> 24
    out = []
    for i in range(n):
        b = (int(word) >> (8 * i)) & 0xFF
        out.append(-((b >> 1) + 1) if b & 1 else (b >> 1) + 1)
    return tuple(out)


def test_pack_normalises_like_the_oracle():
    cnf = CNFInstance(5, [(1, 2, 3), (2, 2, -4), (1, -1, 5), (), (-5,), (3, 3, 3), (4, -2, 4)])
    words, inst = S.pack_batch([cnf])
    assert [_unpack(w) for w in words] == [(1, 2, 3), (2, -4), (), (-5,), (3,), (4, -2)]
    assert inst["num_clauses"][0] == 6 and inst["num_vars"][0] == 5 and inst["clause_begin"][0] == 0


def test_pack_batch_offsets_and_random_instances():
    cnfs = batched_3sat(40)
    words, inst = S.pack_batch(cnfs)
    for i, c in enumerate(cnfs):
        b, n = int(inst["clause_begin"][i]), int(inst["num_clauses"][i])
        assert [_unpack(w) for w in words[b : b + n]] == [tuple(cl) for cl in c.clauses]
    S.validate(words, inst, S.plan_jobs(inst), 6)


def test_limits_raise():
    with pytest.raises(NotImplementedError):
        S.pack_batch([CNFInstance(129, [(1, 2, 129)])])
    with pytest.raises(NotImplementedError):
        S.pack_batch([CNFInstance(4, [(1, 2, 3, 4)])])
    with pytest.raises(NotImplementedError):
        S.pack_batch([CNFInstance(100, [(1 + i % 100, 1 + (i * 7) % 100) for i in range(769)])])
    ok = CNFInstance(128, [(1 + i % 128, -(1 + (i + 1) % 128), 1 + (i + 2) % 128) for i in range(768)])
    words, inst = S.pack_batch([ok])
    assert inst["num_clauses"][0] == 768 and S.supported(ok)


def test_plan_jobs_cover_every_subtree_once():
    cnfs = batched_3sat(200)
    _, inst = S.pack_batch(cnfs)
    for target in (None, 512.0, 64.0, 4.0):
        jobs = S.plan_jobs(inst, split_target=target, max_split=6)
        for i in range(len(cnfs)):
            mine = jobs[jobs["instance"] == i]
            k = int(mine["split"][0])
            assert np.all(mine["split"] == k)
            assert sorted(mine["prefix"].tolist()) == list(range(1 << k))
        if target is None:
            assert np.all(jobs["split"] == 0)
    # hard instances (large n near the threshold) are split more and launched first
    jobs = S.plan_jobs(inst, split_target=16.0, max_split=6)
    est = S.log2_search_size(inst["num_vars"], inst["num_clauses"])
    per_job = est[jobs["instance"]] - jobs["split"]
    assert np.all(np.diff(per_job) <= 1e-9)


def test_validate_rejects_out_of_range_records():
    words, inst = S.pack_batch(batched_3sat(8))
    jobs = S.plan_jobs(inst)
    bad = jobs.copy()
    bad["instance"][0] = 8
    with pytest.raises(ValueError):
        S.validate(words, inst, bad, 6)
    bad = jobs.copy()
    bad["split"][0], bad["prefix"][0] = 2, 4
    with pytest.raises(ValueError):
        S.validate(words, inst, bad, 6)
    bad_inst = inst.copy()
    bad_inst["num_clauses"][-1] += 1
    with pytest.raises(ValueError):
        S.validate(words, bad_inst, jobs, 6)
    small = inst.copy()
    small["num_vars"][0] = 3
    with pytest.raises(ValueError):
        S.validate(words, small, jobs, 6)
    with pytest.raises(ValueError):
        S.validate(words, inst, jobs, 7)


def test_validate_checks_job_start_assignments():
    """Jobs that start below the root (as hand-overs do) carry an assignment: it must name
    only the instance's variables, keep values inside it, and leave the branch stack room
    for every free variable."""
    words, inst = S.pack_batch(batched_3sat(4))
    jobs = S.plan_jobs(inst, split_target=None)
    ok = jobs.copy()
    ok["depth"][0] = 2
    ok["own_from"][0] = 2
    ok["assigned"][0, 0] = 0b11
    ok["value"][0, 0] = 0b01
    S.validate(words, inst, ok, 6)
    bad = ok.copy()
    bad["assigned"][0, 1] = 1  # variable 64 does not exist below 65 variables
    if int(inst["num_vars"][bad["instance"][0]]) <= 64:
        with pytest.raises(ValueError):
            S.validate(words, inst, bad, 6)
    bad = ok.copy()
    bad["value"][0, 0] = 0b100  # a value for an unassigned variable
    with pytest.raises(ValueError):
        S.validate(words, inst, bad, 6)
    bad = ok.copy()
    bad["depth"][0] = 127  # 127 + free variables exceeds the 128-frame stack
    bad["own_from"][0] = 127
    with pytest.raises(ValueError):
        S.validate(words, inst, bad, 6)


def _plan_jobs_numpy(inst, split_target, max_split, tail_fraction=0.0, tail_boost=0):
    """Restatement of the start-job rule in numpy (the C planner's specification)."""
    num = len(inst)
    est = S.log2_search_size(inst["num_vars"], inst["num_clauses"])
    if split_target is None:
        split = np.zeros(num, dtype=np.int64)
    else:
        cap = min(max_split, S.MAX_SPLIT)
        split = np.clip(np.ceil(est - np.log2(split_target)), 0, cap).astype(np.int64)
        if tail_boost > 0 and tail_fraction > 0 and num:
            k = min(num, max(1, int(np.ceil(tail_fraction * num))))
            thr = np.sort(est)[num - k]  # the k-th largest estimate
            split = np.minimum(split + np.where(est >= thr, tail_boost, 0), cap)
    order = np.argsort(-(est - split), kind="stable")
    nj = (np.int64(1) << split)[order]
    jobs = np.zeros(int(nj.sum()), dtype=S.JOB_DTYPE)
    jobs["instance"] = np.repeat(order, nj)
    jobs["split"] = np.repeat(split[order], nj)
    starts = np.zeros(num, dtype=np.int64)
    np.cumsum(nj[:-1], out=starts[1:])
    prefix = np.arange(jobs.size, dtype=np.int64) - np.repeat(starts, nj)
    jobs["prefix"] = prefix
    jobs["own_from"] = np.where(prefix > 0, np.floor(np.log2(np.maximum(prefix, 1))).astype(np.int64) + 1, 0)
    return jobs


def test_plan_jobs_match_the_numpy_rule():
    """The host library's start jobs equal the numpy restatement record for record, over
    clause ratios either side of the threshold and every split setting."""
    from tensorsat_amd.sat import CNFInstance

    cnfs = batched_3sat(300) + [CNFInstance.random(3, n, int(r * n), rng=7 + n) for n in (20, 60, 100)
                               for r in (3.0, 4.26, 5.5)]
    _, inst = S.pack_batch(cnfs)
    for target, ms, tf, tb in ((None, 8, 0.0, 0), (512.0, 8, 0.0, 0), (64.0, 6, 0.0, 0), (4.0, 8, 0.0, 0),
                               (1.0, 3, 0.0, 0), (0.25, 40, 0.0, 0), (8.0, 8, 0.01, 3), (8.0, 8, 0.2, 2),
                               (64.0, 5, 1.0, 4), (8.0, 6, 0.001, 31), (None, 8, 0.5, 3)):
        got = S.plan_jobs(inst, split_target=target, max_split=ms, tail_fraction=tf, tail_boost=tb)
        want = _plan_jobs_numpy(inst, target, ms, tf, tb)
        assert got.tobytes() == want.tobytes(), (target, ms, tf, tb)
        if tb and target is not None and tf < 1.0:  # the tail got more jobs than the rule without it
            assert len(got) > len(_plan_jobs_numpy(inst, target, ms))
    out = np.zeros(4, dtype=S.JOB_DTYPE)  # too small: a new array comes back
    assert S.plan_jobs(inst, split_target=64.0, out=out).tobytes() == _plan_jobs_numpy(inst, 64.0, 8).tobytes()
    with pytest.raises(ValueError):
        S.plan_jobs(inst, split_target=0.0)
    for tf, tb in ((-0.1, 3), (1.5, 3), (0.01, -1), (0.01, 32)):
        with pytest.raises(ValueError):
            S.plan_jobs(inst, split_target=8.0, tail_fraction=tf, tail_boost=tb)
