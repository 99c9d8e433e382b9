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


@pytest.mark.parametrize("width", [8, 12])
def test_graph_replays_match_eager_and_dpll(width):
    import torch

    from tensorsat_amd.solver import BatchSolver

    cnfs = batched_3sat(384)
    want = np.array([dpll.solve(c.num_vars, c.clauses) for c in cnfs], dtype=np.uint8)
    solver = BatchSolver(width=width, first_pass=2)
    solver.prepare(cnfs)
    eager = solver.run()
    np.testing.assert_array_equal(eager, want)
    solver.capture()
    for rep in range(4):
        print(f"width {width} replay {rep}", flush=True)
        solver.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(solver.verdicts(), eager)
        # eager work of the same program between replays
        np.testing.assert_array_equal(solver.run(), eager)
