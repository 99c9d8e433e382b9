"""
GPU parity of exact model counting by the device cube search in exhaustive mode
(``single.count_search``): counts equal brute force on small instances (satisfiable,
unsatisfiable, sparse ones with variables in no clause, no clauses at all), counts of
n = 50..70 instances equal the independent model counter's, the split into jobs and the
hand-over budget do not change a count, and a count beyond 64 bits raises.
"""

import json
from pathlib import Path

import pytest

from oracle import brute
from tensorsat_amd.sat import CNFInstance
from tensorsat_amd.workloads import count_3sat

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_counts_match_brute_force():
    from tensorsat_amd.single import count_search

    cnfs = [CNFInstance.random(3, n, int(r * n), rng=1000 + 7 * n + int(10 * r))
            for n in (8, 12, 16, 20) for r in (1.0, 3.0, 4.3, 6.0)]
    cnfs += [CNFInstance(10, []), CNFInstance(6, [(1, -2), (2, 3, -4), (-1,)]), CNFInstance(5, [(1,), (-1,)])]
    want = [brute.count(c.num_vars, c.clauses) for c in cnfs]
    for split, budget in ((0, None), (3, 4096), (8, 16)):
        r = count_search(cnfs, max_split=split, budget=budget)
        print(f"split {split}: {r.value}", flush=True)
        assert r.value == want


def test_counts_match_independent_counter():
    """n = 50..70 at the #SAT workload's clause ratio (m = 3n), several instances per
    launch, against the independent model counter (``oracle/sharp.c``)."""
    from oracle import sharp
    from tensorsat_amd.single import count_search

    cnfs = [CNFInstance.random(3, n, 3 * n, rng=4000 + n) for n in range(50, 71, 2)]
    want = [sharp.count(c.num_vars, c.clauses) for c in cnfs]
    r = count_search(cnfs)
    print(f"gpu {r.gpu_s:.3f}s, {r.cubes} start jobs, {r.stats['rounds']} rounds", flush=True)
    assert r.value == want


def test_counts_do_not_depend_on_split_or_budget():
    from tensorsat_amd.single import count_search

    cnfs = [CNFInstance.random(3, 60, 180 + 10 * i, rng=77 + i) for i in range(4)]
    a = count_search(cnfs, max_split=2, budget=None).value
    b = count_search(cnfs, max_split=9).value
    c = count_search(cnfs, max_split=4, budget=64).value
    assert a == b == c
    assert any(x > 0 for x in a)


def test_count_beyond_64_bits_raises():
    from tensorsat_amd.single import count_search

    with pytest.raises(OverflowError):
        count_search([CNFInstance(70, [(1, 2, 3)])])
