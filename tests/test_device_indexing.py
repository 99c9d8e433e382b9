"""
CPU emulation of the device step engine's index math (``csrc/device.h``).

Rebuilds the byte-chunk tables exactly as ``build_lut`` does, carrying the tables from
one step to the next the way the persistent kernel keeps them in LDS (entries a step does
not rebuild keep stale values, here filled with large garbage first), and evaluates the
offsets ``step_u8`` computes for every out index and summed index.  Every offset must lie
inside its operand, and the values must match the oracle's einsum on the same records.
"""

import numpy as np
import pytest

from tensorsat_amd import program as PG
from tensorsat_amd.network import bintree_network, clause_network
from tensorsat_amd.sat import CNFInstance

NOPOS = PG.NOPOS


def _build(st, tables, need_c):
    ro, nc = int(st["out_rank"]), int(st["sum_rank"])
    nqo = max(1, (ro + 7) // 8)
    v = np.arange(256)
    for q in range(nqo):
        al = np.zeros(256, np.int64)
        ar = np.zeros(256, np.int64)
        for i in range(8):
            j = q * 8 + i
            if j < ro:
                on = (v >> i) & 1
                pl, pr = int(st["o_lpos"][j]), int(st["o_rpos"][j])
                if pl != NOPOS:
                    al |= on << pl
                if pr != NOPOS:
                    ar |= on << pr
        tables["ol"][q] = al
        tables["or"][q] = ar
    if need_c:
        nqc = max(1, (nc + 7) // 8)
        for q in range(nqc):
            al = np.zeros(256, np.int64)
            ar = np.zeros(256, np.int64)
            for i in range(8):
                j = q * 8 + i
                if j < nc:
                    on = (v >> i) & 1
                    pl, pr = int(st["c_lpos"][j]), int(st["c_rpos"][j])
                    if pl != NOPOS:
                        al |= on << pl
                    if pr != NOPOS:
                        ar |= on << pr
            tables["cl"][q] = al
            tables["cr"][q] = ar


def _offsets(T, idx, nbits, low_from_table=True):
    nq = (nbits + 7) // 8
    r = T[0][idx & 255].copy()
    for q in range(1, 4):
        if nq > q:
            r += T[q][(idx >> (8 * q)) & 255]
    return r


def emulate_job(prog, j, mode):
    job = prog.jobs[j]
    arena = np.zeros(int(job["arena_elems"]), np.uint8)
    lv = prog.job_leaves(j)
    arena[: lv.size] = lv
    rng = np.random.default_rng(j)
    tables = {k: rng.integers(1 << 28, 1 << 30, size=(4, 256)) for k in ("ol", "or", "cl", "cr")}
    for st in prog.job_steps(j):
        ro, nc = int(st["out_rank"]), int(st["sum_rank"])
        rl, rr = int(st["lhs_rank"]), int(st["rhs_rank"])
        norhs = bool(int(st["flags"]) & PG.STEP_NORHS)
        kinner = bool(int(st["flags"]) & PG.STEP_KINNER)
        need_c = (not kinner) and nc > 0
        _build(st, tables, need_c)
        o = np.arange(1 << ro, dtype=np.int64)
        offl = _offsets(tables["ol"], o, ro)
        offr = _offsets(tables["or"], o, ro)
        c = np.arange(1 << nc, dtype=np.int64)
        if kinner:
            cl = c
            cr = c
        elif nc > 0:
            cl = _offsets(tables["cl"], c, nc)
            cr = _offsets(tables["cr"], c, nc)
        else:
            cl = cr = np.zeros(1, np.int64)
        L = offl[:, None] + cl[None, :]
        assert L.min() >= 0 and L.max() < (1 << rl), "lhs offset out of bounds"
        lhs = arena[int(st["lhs_off"]) : int(st["lhs_off"]) + (1 << rl)]
        a = lhs[L].astype(np.int64)
        if not norhs:
            R = offr[:, None] + cr[None, :]
            assert R.min() >= 0 and R.max() < (1 << rr), "rhs offset out of bounds"
            rhs = arena[int(st["rhs_off"]) : int(st["rhs_off"]) + (1 << rr)]
            a = a & rhs[R].astype(np.int64)
        acc = a.sum(axis=1)
        out = ((acc % 256) != 0) if mode == "MOD256" else (acc != 0)
        oo = int(st["out_off"])
        arena[oo : oo + (1 << ro)] = out.astype(np.uint8)
    r0, rr_ = int(job["result_off"]), int(job["result_rank"])
    return arena[r0 : r0 + (1 << rr_)].copy()


@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
def test_batch_program_indexing(mode):
    from oracle import executor as OX

    rng = np.random.default_rng(5)
    cnfs = []
    for i in range(16):
        n = int(rng.integers(14, 21))
        ratio = 4.26 if i % 2 == 0 else 3.5
        cnfs.append(CNFInstance.random(3, n, int(round(ratio * n)), rng=1000 + i))
    prog = PG.sat_batch(cnfs, width=8)
    PG.validate(prog)
    for j in range(len(prog.jobs)):
        got = emulate_job(prog, j, mode)
        ref, _ = OX.run_job(prog, j, mode)
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("enc", [bintree_network, clause_network])
def test_whole_network_indexing(enc):
    from oracle import executor as OX

    for seed in range(4):
        n = 9 + seed
        cnf = CNFInstance.random(3, n, int(round(4.26 * n)), rng=seed)
        d = enc(cnf)
        prog = PG.compile_network(d.wiring, d.blocks, PG.elim(d.wiring, trials=2))
        PG.validate(prog)
        got = emulate_job(prog, 0, "GT0")
        ref, _ = OX.run_job(prog, 0, "GT0")
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("mode", ["GT0", "MOD256"])
def test_lds_arena_programs(mode):
    """Programs that the batched kernel runs with LDS arenas (largest arena <= 64 KiB):
    every tensor then has at most 16 bits, which the two-chunk tables of the LDS variant
    cover, and the table walk (as emulated above) equals the oracle on every job."""
    from oracle import executor as OX
    from tensorsat_amd import _native as N
    from tensorsat_amd.workloads import batched_3sat

    cnfs = batched_3sat(24)
    for width in (8, 12):
        prog = PG.sat_batch(cnfs, width=width, max_cubes=2)
        PG.validate(prog)
        stride = (int(prog.jobs["arena_elems"].max()) + 255) // 256 * 256
        assert stride <= N.BATCH_LDS_MAX
        st = prog.steps
        assert int(max(st["out_rank"].max(), st["sum_rank"].max(), st["lhs_rank"].max(), st["rhs_rank"].max())) <= 16
        for j in range(len(prog.jobs)):
            got = emulate_job(prog, j, mode)
            ref, _ = OX.run_job(prog, j, mode)
            np.testing.assert_array_equal(got, ref)


def _take_positions(num_jobs: int, g: int, seg: int, rng) -> list[int]:
    """Queue positions handed out by the batched kernels' take_job (positions w and
    g + w static, then interleaved segments of the rest, one counter each), with the
    workgroups' takes interleaved in a random order."""
    counters = [0] * seg
    taken = [0] * g
    live = list(range(g))
    out = []
    while live:
        w = live[rng.integers(len(live))]
        got = None
        while taken[w] < 2:
            pos = w if taken[w] == 0 else g + w
            taken[w] += 1
            if pos < num_jobs:
                got = pos
                break
        if got is not None:
            out.append(got)
            continue
        rest, x = num_jobs - 2 * g, w % seg
        seg_len = (rest - x + seg - 1) // seg if rest > x else 0
        if seg_len == 0 or counters[x] >= seg_len:
            live.remove(w)
            continue
        kl = counters[x]
        counters[x] += 1
        if kl < seg_len:
            out.append(2 * g + x + seg * kl)
        else:
            live.remove(w)
    return out


@pytest.mark.parametrize("num_jobs,g", [(0, 8), (5, 8), (8, 8), (9, 8), (16, 8), (17, 8), (100, 8), (100, 13),
                                        (6979, 6144), (12046, 6144), (12046, 8192), (20000, 8192), (50, 64),
                                        (1000, 1000), (2001, 1000), (3, 3), (6, 3)])
def test_job_queue_takes_every_position_once(num_jobs, g):
    from tensorsat_amd import _native as N

    if g < N.BATCH_SEGMENTS and num_jobs > 2 * g:
        pytest.skip("rejected by ts_batch_run")
    rng = np.random.default_rng(num_jobs * 7 + g)
    pos = _take_positions(num_jobs, g, N.BATCH_SEGMENTS, rng)
    assert sorted(pos) == list(range(num_jobs))


@pytest.mark.parametrize("out_wires", [(0, 1), (), (4,), (2, 7, 9)])
def test_permuted_rhs_step_keeps_its_value(out_wires):
    """The general (not K-innermost) one-step programs of the GPU tests: permuting the rhs
    leaf keeps the oracle's value, which equals a direct einsum over the wires."""
    from oracle import executor as OX
    from wide_steps import wide_sum_program

    for seed in range(6):
        for mode, dens in (("GT0", None), ("MOD256", 0.7)):
            plain, leaves = wide_sum_program(seed, out_wires, density=dens, permute=False)
            perm, _ = wide_sum_program(seed, out_wires, density=dens, permute=True)
            assert int(perm.steps["flags"][0]) & PG.STEP_KINNER == 0
            assert int(perm.steps["sum_rank"][0]) == 10 - len(out_wires)
            a = OX.run_job(plain, 0, mode)[0]
            b = OX.run_job(perm, 0, mode)[0]
            np.testing.assert_array_equal(a, b)
            # direct: leaves are in slot wire order, wire i of a slot's list = axis i
            rng = np.random.default_rng(seed)
            wb = [int(x) for x in rng.permutation(10)]
            A = leaves[0].reshape((2,) * 10).astype(np.int64)
            B = leaves[1].reshape((2,) * 10).astype(np.int64)
            letters = "abcdefghij"
            spec = f"{letters},{''.join(letters[x] for x in wb)}->{''.join(letters[x] for x in out_wires)}"
            acc = np.einsum(spec, A, B).reshape(-1)
            want = ((acc % 256) != 0) if mode == "MOD256" else (acc != 0)
            np.testing.assert_array_equal(np.asarray(a).reshape(-1), want.astype(np.uint8))
