"""
The LDS-arena batched kernel (one wave per job) against the device-memory kernel and the
oracle on the same programs: verdicts per instance, with and without early exit, in both
epilogues.
"""

import numpy as np
import pytest

from oracle import executor as OX
from tensorsat_amd import program as PG
from tensorsat_amd.workloads import batched_3sat

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("width", [6, 10, 12])
@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
def test_lds_kernel_matches_global_kernel_and_oracle(width, mode):
    from tensorsat_amd.executor import BatchExecutor

    cnfs = batched_3sat(256, start=1000)
    prog = PG.sat_batch(cnfs, width=width, max_cubes=3)
    print(f"width {width} {mode}: {len(prog.jobs)} jobs, max arena {int(prog.jobs['arena_elems'].max())}", flush=True)
    ref = np.asarray(OX.decide(prog, mode), dtype=np.uint8)
    lds = BatchExecutor(prog, lds=True, track_jobs=True)
    glb = BatchExecutor(prog, lds=False)
    assert lds.lds and not glb.lds
    for early in (False, True):
        a = lds.run(mode, early_exit=early)
        b = glb.run(mode, early_exit=early)
        np.testing.assert_array_equal(a, ref)
        np.testing.assert_array_equal(b, ref)
    # without early exit every job ran exactly once
    lds.run(mode, early_exit=False)
    assert (lds.job_done.cpu().numpy()[: len(prog.jobs)] == 1).all()
