"""
Packet form of compiled programs (``program.packetize`` / ``tsc_packetize``), checked on
the CPU: every job gives the oracle's result of the original program; within a packet no
step reads or writes bytes another step of the packet writes; lanes, masks and packet
spans are laid out as the one-wave kernel reads them; the arena stays within bounds.
"""

import numpy as np
import pytest

from oracle import executor as OX
from tensorsat_amd import program as PG
from tensorsat_amd.workloads import batched_3sat


def _lanes(st) -> int:
    ro, nc, fl = int(st["out_rank"]), int(st["sum_rank"]), int(st["flags"])
    general = not (fl & PG.STEP_KINNER) and nc > 0
    return 1 << max(ro, min(nc, 6) if general else 0)


def _check_layout(prog: PG.Program) -> int:
    npk = 0
    for j in range(len(prog.jobs)):
        st = prog.job_steps(j)
        if len(st) == 0:
            continue
        assert np.all(st["flags"] & PG.STEP_PACKET)
        spans = PG.packet_spans(st)
        assert spans and spans[-1][1] == len(st)
        npk += len(spans)
        for b, e in spans:
            assert e - b <= PG.PACKET_CHUNK
            assert b // PG.PACKET_CHUNK == (e - 1) // PG.PACKET_CHUNK, "packet straddles a staging chunk"
            lane, mask = 0, 0
            writes, reads = [], []
            for k in range(b, e):
                s = st[k]
                assert int(s["pad"][0]) == lane
                mask |= 1 << lane
                lane += _lanes(s)
                o = int(s["out_off"])
                writes.append((o, o + (1 << int(s["out_rank"]))))
                lo = int(s["lhs_off"])
                reads.append((lo, lo + (1 << int(s["lhs_rank"]))))
                if not int(s["flags"]) & PG.STEP_NORHS:
                    ro_ = int(s["rhs_off"])
                    reads.append((ro_, ro_ + (1 << int(s["rhs_rank"]))))
            assert lane <= PG.PACKET_LANES
            got = int.from_bytes(bytes(st[b]["o_lpos"][16:24]), "little")
            assert got == mask
            for i, (a0, a1) in enumerate(writes):
                for k, (b0, b1) in enumerate(writes):
                    if k != i:
                        assert a1 <= b0 or b1 <= a0, "two writes of one packet overlap"
                for (b0, b1) in reads:
                    assert a1 <= b0 or b1 <= a0, "a packet reads what it writes"
    return npk


@pytest.mark.parametrize("width,mode", [(6, "GT0"), (6, "MOD256"), (5, "GT0")])
def test_packetized_batch_matches_original(width, mode):
    cnfs = batched_3sat(96, start=500)
    prog = PG.sat_batch(cnfs, width=width, max_cubes=3)
    pk = PG.packetize(prog)
    PG.validate(pk)
    npk = _check_layout(pk)
    assert npk < len(prog.steps)  # fewer packet rounds than steps
    for j in range(len(prog.jobs)):
        a, _ = OX.run_job(prog, j, mode)
        b, _ = OX.run_job(pk, j, mode)
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(OX.decide(prog, mode), OX.decide(pk, mode))


def test_packetized_wide_steps_and_counts():
    """Non-K-innermost steps with more summed bits than out bits (lanes from the summed
    bits) and the counting semiring on the same packet form."""
    from wide_steps import wide_sum_program

    for seed in range(4):
        for ow in ((0, 1), (), (2, 7, 9)):
            prog, _ = wide_sum_program(seed, ow)
            pk = PG.packetize(prog)
            PG.validate(pk)
            _check_layout(pk)
            for mode in ("GT0", "MOD256", "COUNT"):
                np.testing.assert_array_equal(OX.run_job(prog, 0, mode)[0], OX.run_job(pk, 0, mode)[0])


def test_packetize_rejects_wide_outputs_and_repeats():
    from wide_steps import wide_sum_program

    prog, _ = wide_sum_program(0, (0, 1, 2, 3, 4, 5, 6))
    with pytest.raises(ValueError):
        PG.packetize(prog)
    cnfs = batched_3sat(8, start=40)
    pk = PG.packetize(PG.sat_batch(cnfs, width=6, max_cubes=2))
    with pytest.raises(ValueError):
        PG.packetize(pk)


def _run_lanes(lp: PG.LaneProgram, j: int, mode: str) -> np.ndarray:
    """CPU interpreter of one lane-form job, following the kernel: packets in order, each
    lane from its own record, stores after the whole packet is read."""
    job = lp.jobs[j]
    arena = np.zeros(int(job["arena_elems"]), np.int64)
    lv = lp.source.job_leaves(j)
    arena[: lv.size] = lv
    recs = lp.job_lanes(j)
    q = 0
    while q < len(recs):
        head = recs[q]
        cnt = max(int(head["count"]), 1)
        if int(head["op"]) & PG.LANE_SKIP:
            assert (q + cnt) % PG.LANE_CHUNK == 0
            q += cnt
            continue
        assert q // PG.LANE_CHUNK == (q + cnt - 1) // PG.LANE_CHUNK and cnt <= 64
        pk = recs[q : q + cnt]
        stores = []
        for r in pk:
            op, nc = int(r["op"]), int(r["op"]) & 7
            norhs = bool(op & PG.LANE_NORHS)
            lhs, rhs = int(r["lhs"]), int(r["rhs"])
            if op & PG.LANE_GENERAL:
                first = int(r["first"])
                acc = 0
                for c in range(1 << nc):
                    t = pk[first + c]
                    a = arena[lhs + int(t["tl"])]
                    acc += a & (1 if norhs else arena[rhs + int(t["tr"])])
            else:
                a = arena[lhs : lhs + (1 << nc)]
                acc = int((a & (1 if norhs else arena[rhs : rhs + (1 << nc)])).sum())
            if int(r["out"]) != PG.LANE_NONE:
                stores.append((int(r["out"]), int((acc % 256) != 0) if mode == "MOD256" else int(acc != 0)))
        for o, v in stores:
            arena[o] = v
        q += cnt
    r0, rr = int(job["result_off"]), int(job["result_rank"])
    return arena[r0 : r0 + (1 << rr)].astype(np.uint8)


@pytest.mark.parametrize("width,mode", [(6, "GT0"), (6, "MOD256"), (4, "GT0")])
def test_lane_form_matches_oracle(width, mode):
    cnfs = batched_3sat(64, start=700)
    prog = PG.sat_batch(cnfs, width=width, max_cubes=3)
    lp = PG.lanes(PG.packetize(prog))
    PG.validate_lanes(lp)
    assert np.all(lp.jobs["result_off"] == lp.source.jobs["result_off"])
    for j in range(len(prog.jobs)):
        np.testing.assert_array_equal(_run_lanes(lp, j, mode), OX.run_job(prog, j, mode)[0])


def test_lane_form_general_steps():
    from wide_steps import wide_sum_program

    for seed in range(4):
        for ow in ((0, 1), (4, 5), (1, 3, 5), (0, 2, 4, 6)):
            prog, _ = wide_sum_program(seed, ow, n=8)
            lp = PG.lanes(PG.packetize(prog))
            PG.validate_lanes(lp)
            assert np.any(lp.lanes["op"] & PG.LANE_GENERAL)
            for mode in ("GT0", "MOD256"):
                np.testing.assert_array_equal(_run_lanes(lp, 0, mode), OX.run_job(prog, 0, mode)[0])


def test_lane_form_rejects_wide_sums():
    from wide_steps import wide_sum_program

    prog, _ = wide_sum_program(0, (4,), n=8)  # 7 summed bits
    with pytest.raises(ValueError):
        PG.lanes(PG.packetize(prog))


def test_lds_variant_rejects_wide_summed_steps():
    """The LDS table engine covers 16 summed bits: a step with more than 6 out bits and
    more than 16 summed bits must not run there (it goes to device arenas)."""
    from tensorsat_amd.executor import lds_fits

    jobs = np.zeros(1, dtype=PG.JOB_DTYPE)
    jobs["num_steps"] = 1
    jobs["arena_elems"] = 4096
    steps = np.zeros(1, dtype=PG.STEP_DTYPE)
    steps["out_rank"], steps["sum_rank"] = 7, 16
    prog = PG.Program(jobs, steps, np.zeros(16, np.uint8), np.zeros((1, PG.NSTAT), np.int32))
    assert lds_fits(prog) == (True, "")
    steps["sum_rank"] = 17
    ok, why = lds_fits(prog)
    assert not ok and "16 bits" in why
    steps["out_rank"] = 6  # one-wave steps are not built on the tables
    assert lds_fits(prog)[0]
    jobs["arena_elems"] = 70000
    assert not lds_fits(prog)[0]
