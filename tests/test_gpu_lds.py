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


@pytest.mark.parametrize("packets", [False, True])
@pytest.mark.parametrize("out_wires", [(0, 1), (), (4,), (2, 7, 9)])
@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
def test_one_wave_kernel_general_steps_with_many_summed_bits(out_wires, mode, packets):
    """A general step (summed bits not innermost) summing 7 to 10 bits: the one-wave
    kernel's block loop over summed bits past the sixth, against the oracle."""
    from tensorsat_amd.executor import BatchExecutor
    from wide_steps import wide_sum_program

    seen = set()
    for seed in range(12):
        prog, _ = wide_sum_program(seed, out_wires, density=0.7 if mode == "MOD256" else None)
        want = np.uint8(np.asarray(OX.run_job(prog, 0, mode)[0]).any())
        seen.add(int(want))
        ex = BatchExecutor(prog, lds=True, packets=packets)
        assert ex.one_wave and ex.packets == packets
        assert ex.lane_form == (packets and int(prog.steps["sum_rank"].max()) <= 6)
        glb = BatchExecutor(prog, lds=False)
        print(f"{mode} out {out_wires} seed {seed} packets {packets}", flush=True)
        assert ex.run(mode, early_exit=False)[0] == want
        assert glb.run(mode, early_exit=False)[0] == want
    if mode == "GT0":
        assert seen == {0, 1}


@pytest.mark.parametrize("width", [4, 5, 6])
@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
def test_packet_kernel_matches_step_kernel_and_oracle(width, mode):
    """One-wave kernel in packet form (independent steps of a job sharing the wave's
    lanes) against the same kernel step by step and the oracle, with and without early
    exit; without it every job runs once."""
    from tensorsat_amd.executor import BatchExecutor

    cnfs = batched_3sat(512, start=2000)
    prog = PG.sat_batch(cnfs, width=width, max_cubes=4)
    ref = np.asarray(OX.decide(prog, mode), dtype=np.uint8)
    ln = BatchExecutor(prog, track_jobs=True)
    pk = BatchExecutor(prog, track_jobs=True, lanes=False)
    st = BatchExecutor(prog, packets=False)
    assert ln.lane_form and pk.packets and not pk.lane_form and pk.one_wave and not st.packets
    print(f"width {width} {mode}: {len(prog.jobs)} jobs, {len(prog.steps)} steps", flush=True)
    for early in (False, True):
        np.testing.assert_array_equal(ln.run(mode, early_exit=early), ref)
        np.testing.assert_array_equal(pk.run(mode, early_exit=early), ref)
        np.testing.assert_array_equal(st.run(mode, early_exit=early), ref)
    for ex in (ln, pk):
        ex.run(mode, early_exit=False)
        assert (ex.job_done.cpu().numpy()[: len(prog.jobs)] == 1).all()


@pytest.mark.parametrize("out_wires", [(0, 1), (4, 5), (1, 3, 5), (0, 2, 4, 6)])
@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
def test_lane_form_general_steps(out_wires, mode):
    """General steps (summed bits not innermost, up to 6 of them) in lane form: the summed
    terms read from the table lanes of each step."""
    from tensorsat_amd.executor import BatchExecutor
    from wide_steps import wide_sum_program

    seen = set()
    for seed in range(12):
        prog, _ = wide_sum_program(seed, out_wires, n=8, density=0.7 if mode == "MOD256" else 0.08)
        want = np.uint8(np.asarray(OX.run_job(prog, 0, mode)[0]).any())
        seen.add(int(want))
        ex = BatchExecutor(prog, lds=True)
        assert ex.lane_form
        print(f"{mode} out {out_wires} seed {seed}", flush=True)
        assert ex.run(mode, early_exit=False)[0] == want
    if mode == "GT0":
        assert seen == {0, 1}
