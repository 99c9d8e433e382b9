"""
GPU parity for the single-instance workloads: the n=120 decision (cubes up to 2^26) and
exact #SAT counts, against the oracle, brute force, DPLL and decomposition invariance.
"""

import numpy as np
import pytest

from oracle import brute, dpll
from oracle import executor as OX
from tensorsat_amd import program as PG
from tensorsat_amd.sat import CNFInstance
from tensorsat_amd.workloads import count_3sat, single_3sat

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
def test_n120_cube_intermediates_bit_exact(mode):
    """Every intermediate of the first cubes of the n=120 instance (width 14 so the oracle
    keeps up) equals the oracle's, byte for byte."""
    from tensorsat_amd.executor import DeviceProgram, run_job

    cnf = single_3sat(0)
    prog = PG.sat_batch([cnf], width=14, max_cubes=3, plan_trials=2)
    dp = DeviceProgram(prog)
    for j in range(len(prog.jobs)):
        print(f"cube {j} steps={int(prog.jobs['num_steps'][j])}", flush=True)
        res, inter = run_job(dp, j, mode, keep_intermediates=True)
        ref, ref_inter = OX.run_job(prog, j, mode, keep_intermediates=True)
        for s, (a, b) in enumerate(zip(inter, ref_inter)):
            np.testing.assert_array_equal(a.cpu().numpy(), b, err_msg=f"cube {j} step {s}")
        np.testing.assert_array_equal(res.cpu().numpy(), ref)


def test_n120_decisions_match_dpll():
    from tensorsat_amd.single import decide

    for seed in range(3):
        cnf = single_3sat(seed)
        r = decide(cnf, width=26)
        print(f"seed {seed}: verdict {r.value} cubes {r.cubes_run}/{r.cubes} gpu {r.gpu_s:.2f}s", flush=True)
        assert r.value == int(dpll.solve(cnf.num_vars, cnf.clauses))
        assert r.stats["max_rank"] <= 26


def test_counts_match_brute_force_and_oracle():
    from tensorsat_amd.executor import BatchCounter

    cnfs = [CNFInstance.random(3, n, 3 * n, rng=300 + n) for n in range(14, 25, 2)]
    prog = PG.sat_batch(cnfs, width=8)
    got = BatchCounter(prog).run()
    assert got == [brute.count(c.num_vars, c.clauses) for c in cnfs]
    assert got == OX.count(prog)


def test_n100_count_is_decomposition_invariant():
    """#SAT of the n=100, m=300 instance: the same exact count from two different cube
    decompositions (widths 14 and 18) -- any dropped, doubled or wrapped term would
    change one of them."""
    from tensorsat_amd.single import count

    cnf = count_3sat(1)
    a = count(cnf, width=14)
    b = count(cnf, width=18)
    print(f"count {a.value} (cubes {a.cubes} / {b.cubes}, gpu {a.gpu_s:.2f}s / {b.gpu_s:.2f}s)", flush=True)
    assert a.value == b.value
    assert a.value > 0 and dpll.solve(cnf.num_vars, cnf.clauses)


def test_count_overflow_raises():
    """All-true network over 70 free variables: 2^70 models do not fit uint64."""
    from tensorsat_amd.executor import BatchCounter

    cnf = CNFInstance(70, [(1, -1)])  # a tautology only: every assignment is a model
    prog = PG.sat_batch([cnf], width=8)
    with pytest.raises(OverflowError):
        BatchCounter(prog).run()


@pytest.mark.parametrize("width,small_bits", [(14, 10), (26, 14), (20, 27)])
def test_fused_runner_leaves_the_same_arena(width, small_bits):
    """ts_program_run_fused (runs of small steps on one workgroup, large steps grid-wide)
    executes the same steps in the same order as the per-step runner: the whole final
    arena of every cube is byte-identical, in both modes."""
    import ctypes

    import torch

    from tensorsat_amd import _native as N
    from tensorsat_amd.executor import DeviceProgram, run_job_fused

    prog = PG.sat_batch([single_3sat(1)], width=width, max_cubes=4, plan_trials=2, threads=8)
    dp = DeviceProgram(prog)
    L = N.lib()
    for mode in ("GT0", "MOD256"):
        for j in range(len(prog.jobs)):
            size = max(16, int(prog.jobs["arena_elems"][j]))
            a = torch.zeros(size, dtype=torch.uint8, device="cuda")
            b = torch.zeros(size, dtype=torch.uint8, device="cuda")
            N.check(L.ts_program_run(dp.job_ptr(j), N.host_ptr(dp.steps_host), N.dptr(dp.leaves), N.dptr(a),
                                     N.MODES[mode], None, None, N.stream_ptr()), "ts_program_run")
            res = run_job_fused(dp, j, mode, arena=b, small_bits=small_bits)
            torch.cuda.synchronize()
            print(f"width {width} small {small_bits} {mode} cube {j}: result {int(res.item())}", flush=True)
            assert torch.equal(a, b), f"cube {j}"
