"""
The batched decision captured as one HIP graph: every replay (with eager launches of the
same passes and host synchronisations between replays) gives the eager verdicts, which
equal the independent DPLL's.
"""

import numpy as np
import pytest

from oracle import dpll
from tensorsat_amd.workloads import batched_3sat

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("merge", [True, False])
@pytest.mark.parametrize("width", [6, 8, 12])
def test_graph_replays_match_eager_and_dpll(width, merge):
    import torch

    from tensorsat_amd.solver import BatchSolver

    cnfs = batched_3sat(384)
    want = np.array([dpll.solve(c.num_vars, c.clauses) for c in cnfs], dtype=np.uint8)
    solver = BatchSolver(width=width, first_pass=2, merge_passes=merge)
    solver.prepare(cnfs)
    print(f"width {width} merge {merge}: {solver.stats.pass2_instances} second-pass instances, "
          f"{len(solver.execs)} launches", flush=True)
    assert len(solver.execs) == (1 if merge or not solver.stats.pass2_jobs else 2)
    eager = solver.run()
    np.testing.assert_array_equal(eager, want)
    solver.capture(steps=1 if width != 6 else 3)  # several decisions in one graph too
    for rep in range(4):
        print(f"width {width} replay {rep}", flush=True)
        solver.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(solver.verdicts(), eager)
        # eager work of the same program between replays
        np.testing.assert_array_equal(solver.run(), eager)
