"""
The batch autotuners on the GPU: the cube-width autotuner of the cube executor
(``solver.autotune_width``) and the search-configuration autotuner of the device search
(``search.autotune``).  Each must return one of its candidates, the one its own table
rates fastest, with a table entry per candidate; and the configuration it picks must
decide the batch with the verdicts of the independent DPLL.
"""

import numpy as np
import pytest

from oracle import dpll
from tensorsat_amd.workloads import batched_3sat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    cnfs = batched_3sat(512)
    want = np.array([dpll.solve(c.num_vars, c.clauses) for c in cnfs], dtype=np.int64)
    return cnfs, want


def test_cube_width_autotuner(batch):
    from tensorsat_amd.solver import BatchSolver, autotune_width

    cnfs, want = batch
    w, table = autotune_width(cnfs, widths=(6, 8), sample=256, reps=2)
    print({k: round(v["cost_s"], 4) for k, v in table.items()}, "->", w, flush=True)
    assert set(table) == {6, 8} and w in table
    assert table[w]["cost_s"] == min(r["cost_s"] for r in table.values())
    for r in table.values():
        assert r["plan_s"] > 0 and r["gpu_ms"] > 0 and r["jobs"] > 0
    s = BatchSolver(width=w)
    s.prepare(cnfs)
    np.testing.assert_array_equal(np.asarray(s.run(), dtype=np.int64), want)


def test_search_autotuner(batch):
    from tensorsat_amd.search import SearchSolver, autotune

    cnfs, want = batch
    choice, table = autotune(cnfs, widths=(4, 6), split_targets=(None, 8.0), budgets=(None, 32), reps=2)
    print(choice, {k: round(v["ms"], 3) for k, v in table.items()}, flush=True)
    assert len(table) == 2 * (1 + 2) * 2  # the tail boost (3 or 0) only with a split target
    best = min(table.values(), key=lambda r: r["ms"])
    assert (choice["width"], choice["split_target"], choice["tail_boost"], choice["budget"]) == \
        (best["width"], best["split_target"], best["tail_boost"], best["budget"])
    for r in table.values():
        assert r["ms"] > 0 and r["rounds"] >= 1
        if r["budget"] is None:
            assert r["rounds"] == 1
    s = SearchSolver(**choice)
    np.testing.assert_array_equal(s.run(cnfs), want)
