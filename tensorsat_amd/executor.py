"""
Device executors for compiled programs (``tensorsat_amd.program``).

* :func:`run_job` / :func:`count_job` — one job, every step a grid-wide kernel on the
  current stream (single large instance: the n=120 decision and the #SAT counts).
* :class:`BatchExecutor` — all jobs of a program in one persistent launch
  (``ts_batch_run``), one workgroup per job at a time, verdicts OR-ed per instance.

These replace the reference's per-step loop (``contractions/cotengra.py:120-126``,
``contractions/simple.py:251-292``) and final ``bool()`` (``diagrams.py:949-958``).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from . import program as PG
from .program import JOB_DTYPE, Program, validate


def _torch():
    import torch

    return torch


class DeviceProgram:
    """Device copies of a program's step, job and leaf arrays."""

    def __init__(self, prog, device: str = "cuda") -> None:
        """``prog``: a Program, or a LaneProgram (its lane records then take the place of
        the step records)."""
        torch = _torch()
        N.lib()
        lane_form = isinstance(prog, PG.LaneProgram)
        if lane_form:
            PG.validate_lanes(prog)
        else:
            validate(prog)
        self.prog = prog
        self.device = device
        self.steps_host = np.ascontiguousarray(prog.lanes if lane_form else prog.steps)
        self.jobs_host = np.ascontiguousarray(prog.jobs)
        leaves = prog.leaves if prog.leaves.size else np.zeros(16, np.uint8)
        self.leaves = torch.from_numpy(np.ascontiguousarray(leaves)).to(device)
        steps_bytes = self.steps_host.view(np.uint8) if self.steps_host.size else np.zeros(128, np.uint8)
        self.steps = torch.from_numpy(np.ascontiguousarray(steps_bytes)).to(device)
        jobs_bytes = self.jobs_host.view(np.uint8) if self.jobs_host.size else np.zeros(JOB_DTYPE.itemsize, np.uint8)
        self.jobs = torch.from_numpy(np.ascontiguousarray(jobs_bytes)).to(device)

    def job_ptr(self, j: int) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.jobs_host.ctypes.data + j * JOB_DTYPE.itemsize)


def run_job(dp: DeviceProgram, j: int, mode: str = "GT0", *, keep_intermediates: bool = False):
    """Run job ``j`` grid-wide; returns (result uint8 tensor, list of step outputs)."""
    torch = _torch()
    L = N.lib()
    job = dp.jobs_host[j]
    arena = torch.empty(max(16, int(job["arena_elems"])), dtype=torch.uint8, device=dp.device)
    inter = None
    offsets = None
    sizes: list[int] = []
    if keep_intermediates:
        steps = dp.prog.job_steps(j)
        sizes = [1 << int(r) for r in steps["out_rank"]]
        offsets = np.zeros(max(1, len(sizes)), dtype=np.int64)
        if sizes:
            offsets[1:] = np.cumsum(sizes)[:-1] if len(sizes) > 1 else []
        inter = torch.empty(max(16, sum(sizes)), dtype=torch.uint8, device=dp.device)
    N.check(
        L.ts_program_run(
            dp.job_ptr(j), N.host_ptr(dp.steps_host), N.dptr(dp.leaves), N.dptr(arena), N.MODES[mode],
            N.dptr(inter) if inter is not None else None,
            N.host_ptr(offsets) if offsets is not None else None, N.stream_ptr(),
        ),
        "ts_program_run",
    )
    r0, rr = int(job["result_off"]), int(job["result_rank"])
    result = arena[r0 : r0 + (1 << rr)].clone()
    outs = []
    if inter is not None:
        for off, sz in zip(offsets.tolist(), sizes):
            outs.append(inter[off : off + sz])
    return result, outs


def run_job_fused(dp: DeviceProgram, j: int, mode: str = "GT0", *, arena=None, small_bits: int = 12):
    """Run job ``j`` on the current stream with ``ts_program_run_fused`` (runs of steps of
    at most ``small_bits`` bits on one workgroup per run, larger steps grid-wide);
    returns a view of the result in the arena (``arena``: a uint8 tensor of at least the
    job's arena size, else a new one)."""
    torch = _torch()
    L = N.lib()
    job = dp.jobs_host[j]
    need = max(16, int(job["arena_elems"]))
    if arena is None:
        arena = torch.empty(need, dtype=torch.uint8, device=dp.device)
    elif arena.numel() < need or arena.dtype != torch.uint8:
        raise ValueError("arena too small for the job")
    N.check(
        L.ts_program_run_fused(dp.job_ptr(j), N.host_ptr(dp.steps_host), N.dptr(dp.steps), N.dptr(dp.leaves),
                               N.dptr(arena), N.MODES[mode], small_bits, N.stream_ptr()),
        "ts_program_run_fused",
    )
    r0, rr = int(job["result_off"]), int(job["result_rank"])
    return arena[r0 : r0 + (1 << rr)]


def gang_stride(dp: DeviceProgram) -> int:
    """Bytes per job arena in a gang: the largest job arena (rounded up to 256) plus room
    for the bit masks of the rows of its largest packable step (summed bits innermost in
    both operands, 4..10 of them, at least 2^12 outputs, at least two out bits held by one
    operand only: ``ts_program_run_gang`` packs those rows when the stride has room)."""
    cached = getattr(dp, "_gang_stride", None)
    if cached is not None:
        return cached
    jobs, st = dp.jobs_host, dp.steps_host
    base = (np.maximum(16, jobs["arena_elems"].astype(np.int64)) + 255) & ~255
    extra = np.zeros(len(jobs), dtype=np.int64)
    if len(st):
        ro = st["out_rank"].astype(np.int64)
        k = st["sum_rank"].astype(np.int64)
        bit = np.arange(PG.MAX_BITS)[None, :] < ro[:, None]
        one_side = (bit & ((st["o_lpos"] == PG.NOPOS) != (st["o_rpos"] == PG.NOPOS))).sum(axis=1)
        ok = (((st["flags"] & PG.STEP_KINNER) != 0) & ((st["flags"] & PG.STEP_NORHS) == 0) & (k >= 4) & (k <= 10)
              & (ro >= 12) & (one_side >= 2))
        need = np.where(ok, ((np.left_shift(1, np.maximum(st["lhs_rank"].astype(np.int64) - 3, 0)) + 15) & ~15)
                        + np.left_shift(1, np.maximum(st["rhs_rank"].astype(np.int64) - 3, 0)), 0)
        for j in range(len(jobs)):  # once per program (cached on it)
            b, n = int(jobs["step_begin"][j]), int(jobs["num_steps"][j])
            if n:
                extra[j] = int(need[b : b + n].max())
    dp._gang_stride = int((int((base + extra).max()) + 255) & ~255)
    return dp._gang_stride


def run_gang(dp: DeviceProgram, first: int, count: int, mode: str = "GT0", *, arena=None, results=None,
             small_bits: int = 12):
    """Run jobs ``first .. first+count-1`` together with ``ts_program_run_gang`` on the
    current stream; returns a uint8 tensor of ``count`` flags (1: the job's result has a
    nonzero entry).  ``arena``: uint8 tensor of at least ``count * gang_stride(dp)``
    bytes, else a new one; ``results``: uint8 tensor of at least ``count`` entries."""
    torch = _torch()
    L = N.lib()
    if first < 0 or count < 0 or first + count > len(dp.jobs_host):
        raise ValueError("job range out of the program")
    stride = gang_stride(dp)
    if arena is None:
        arena = torch.empty(max(1, count) * stride, dtype=torch.uint8, device=dp.device)
    elif arena.numel() < count * stride or arena.dtype != torch.uint8:
        raise ValueError("arena too small for the gang")
    if results is None:
        results = torch.empty(max(1, count), dtype=torch.uint8, device=dp.device)
    elif results.numel() < count or results.dtype != torch.uint8:
        raise ValueError("results too small for the gang")
    N.check(
        L.ts_program_run_gang(dp.job_ptr(first), count, N.host_ptr(dp.steps_host), N.dptr(dp.steps),
                              N.dptr(dp.leaves), N.dptr(arena), stride, N.dptr(results), N.MODES[mode], small_bits,
                              N.stream_ptr()),
        "ts_program_run_gang",
    )
    return results[:count]


def count_job(dp: DeviceProgram, j: int, acc=None, overflow=None):
    """Exact uint64 count of job ``j`` (times 2^free_vars) added into ``acc``."""
    torch = _torch()
    L = N.lib()
    job = dp.jobs_host[j]
    arena = torch.empty(max(2, int(job["arena_elems"])), dtype=torch.int64, device=dp.device)
    if acc is None:
        acc = torch.zeros(1, dtype=torch.int64, device=dp.device)
    if overflow is None:
        overflow = torch.zeros(1, dtype=torch.int32, device=dp.device)
    N.check(
        L.ts_program_count(dp.job_ptr(j), N.host_ptr(dp.steps_host), N.dptr(dp.leaves), N.dptr(arena),
                           N.dptr(acc), N.dptr(overflow), N.stream_ptr()),
        "ts_program_count",
    )
    return acc, overflow


def count_program(dp: DeviceProgram) -> list[int]:
    """Per-instance exact model counts; raises OverflowError if any count >= 2^64."""
    torch = _torch()
    n = dp.prog.num_instances
    accs = torch.zeros(n, dtype=torch.int64, device=dp.device)
    ovf = torch.zeros(n, dtype=torch.int32, device=dp.device)
    for j in range(len(dp.jobs_host)):
        i = int(dp.jobs_host["instance"][j])
        count_job(dp, j, accs[i : i + 1], ovf[i : i + 1])
    torch.cuda.synchronize()
    if int(ovf.max().item()) if n else 0:
        raise OverflowError("model count exceeds 2^64")
    vals = accs.cpu().numpy().view(np.uint64)
    return [int(v) for v in vals]


def job_work(prog: Program) -> np.ndarray:
    """Latency estimate per job: steps run one after the other, and a step takes one pass
    of the workgroup per 256 out x summed entries."""
    jobs = prog.jobs
    steps = prog.steps
    work = np.zeros(len(jobs), dtype=np.float64)
    if len(jobs) and len(steps):
        per = 1.0 + np.ldexp(1.0, steps["out_rank"].astype(np.int64) + steps["sum_rank"].astype(np.int64)) / 256.0
        csum = np.concatenate([[0.0], np.cumsum(per)])
        b = jobs["step_begin"].astype(np.int64)
        e = b + jobs["num_steps"].astype(np.int64)
        work = csum[e] - csum[b]
    return work


def dispatch_order(prog: Program, num_workgroups: int | None = None) -> np.ndarray:
    """
    Job order for the persistent queue: by cube index within the instance (so early
    exit can skip later cubes of decided instances), then longest job first.

    With ``num_workgroups`` = G, the two statically assigned rounds (workgroup w takes
    positions w and G + w) keep their members but are reordered: the m second-round jobs
    go, longest first, to workgroups 0 .. m-1 (dispatched first), whose first jobs are the
    m shortest of the first round, shortest first (long second jobs pair with short first
    ones); the other first-round jobs follow longest first, so the workgroups dispatched
    last get the shortest of them.
    """
    jobs = prog.jobs
    if len(jobs) == 0:
        return np.zeros(0, dtype=np.int32)
    work = job_work(prog)
    order = np.lexsort((-work, jobs["cube"])).astype(np.int32)
    if num_workgroups:
        g = int(num_workgroups)
        r1, r2 = order[:g], order[g : 2 * g]
        m = len(r2)
        r1 = r1[np.argsort(work[r1], kind="stable")]  # shortest first
        rest = r1[m:]
        order = np.concatenate([r1[:m], rest[np.argsort(-work[rest], kind="stable")],
                                r2[np.argsort(-work[r2], kind="stable")], order[2 * g :]]).astype(np.int32)
    return order


ARENA_BUDGET = 16 << 30
"""Device memory for the per-workgroup arenas of the device-memory batched kernels (bytes):
programs with large cubes run on fewer workgroups rather than allocating beyond it."""

ROUND_BOUNDS = (0, 1 << 30)
"""Cube-index boundaries of the dispatch rounds.  Default: one round per program; jobs
of decided instances are skipped by the per-job check inside the persistent kernel.
Finer boundaries such as (0, 1, 2, 4, 8, 16, 64, 1 << 30) add a compaction pass between
rounds but each round then waits for its longest job (measured slower on the batched
workload: 0.61 vs 0.41 ms per 4096 instances at W=10)."""


def lds_fits(prog: Program) -> tuple[bool, str]:
    """Whether the LDS-arena variants of the batched kernel can run ``prog``: every job
    arena within ``BATCH_LDS_MAX`` bytes, and every step wider than one wave (more than
    6 out bits, which runs on the 4-wave table engine) summing at most 16 bits -- that
    engine keeps two 8-bit chunk tables per operand in LDS, so higher summed bits would
    never be visited.  Returns (fits, reason if not)."""
    nj = len(prog.jobs)
    need = (int(prog.jobs["arena_elems"].max()) if nj else 16) + 255 & ~255
    if need > N.BATCH_LDS_MAX:
        return False, f"job arenas of {need} bytes exceed the LDS limit"
    st = prog.steps
    if nj and len(st):
        wide = st["out_rank"] > 6
        if np.any(wide) and int(st["sum_rank"][wide].max()) > 16:
            return False, "a step with more than 6 out bits sums more than 16 bits (LDS table engine limit)"
    return True, ""


class BatchExecutor:
    """
    All jobs of a program in a few persistent launches (``ts_batch_run``): one round per
    range of cube indices; from the second round on, ``ts_batch_compact`` first drops the
    jobs of instances already decided SAT, on the device.
    """

    def __init__(self, prog: Program, *, num_workgroups: int | None = None, device: str = "cuda",
                 order: np.ndarray | None = None, track_jobs: bool = False,
                 round_bounds=ROUND_BOUNDS, lds: bool | None = None, packets: bool | None = None,
                 lanes: bool | None = None) -> None:
        """``lds``: keep each job's arena in LDS (the LDS variant of the batched kernel);
        default: whenever the largest arena fits (``BATCH_LDS_MAX`` bytes).  ``packets``:
        run one-wave programs in packet form (``program.packetize``: independent steps of
        a job share the wave's lanes); default: whenever the program runs on one-wave
        workgroups.  ``lanes``: run the packet form as lane records (``program.lanes``:
        all index math on the host); default: whenever the packet form runs and every step
        sums at most 6 bits."""
        torch = _torch()
        L = N.lib()
        self.prog = prog
        nj = len(prog.jobs)

        def stride(p: Program) -> int:
            return (int(p.jobs["arena_elems"].max()) if nj else 16) + 255 & ~255

        fits, why = lds_fits(prog)
        self.lds = fits if lds is None else bool(lds)
        if self.lds and not fits:
            raise ValueError(why)
        # Every tensor of at most 6 bits: one-wave workgroups (every step <= 64 entries).
        st = prog.steps
        small = bool(nj) and self.lds and (len(st) == 0 or int(st["out_rank"].max()) <= 6)
        self.one_wave = small
        self.packets = small if packets is None else bool(packets) and small
        run = PG.packetize(prog) if self.packets and len(st) else prog
        if self.packets and stride(run) > N.BATCH_LDS_MAX:
            self.packets, run = False, prog  # packet lifetimes need more arena than LDS holds
        self.run_prog = run
        can_lanes = self.packets and len(st) > 0 and int(st["sum_rank"].max()) <= 6 and stride(run) <= 0xFFFF
        self.lane_form = can_lanes if lanes is None else bool(lanes) and can_lanes
        self.dp = DeviceProgram(PG.lanes(run) if self.lane_form else run, device)
        self.arena_stride = stride(run)
        if num_workgroups is None:
            mx = N.i32(0)
            if self.lds:
                variant = 3 if self.lane_form else 2 if self.packets else 1 if self.one_wave else 0
                N.check(L.ts_batch_max_workgroups_lds(self.arena_stride, variant, ctypes.byref(mx)),
                        "ts_batch_max_workgroups_lds")
            else:
                N.check(L.ts_batch_max_workgroups(ctypes.byref(mx)), "ts_batch_max_workgroups")
            num_workgroups = int(mx.value)
        if not self.lds:  # device arenas within budget
            num_workgroups = min(int(num_workgroups), max(N.BATCH_SEGMENTS, ARENA_BUDGET // self.arena_stride))
        # at least one workgroup per non-empty queue segment
        self.num_workgroups = max(1, min(max(int(num_workgroups), N.BATCH_SEGMENTS), max(1, nj)))
        n_arena = 16 if self.lds else self.num_workgroups * self.arena_stride
        self.arenas = torch.empty(n_arena, dtype=torch.uint8, device=device)
        one_round = order is not None or (len(round_bounds) == 2 and round_bounds[0] <= 0
                                          and (nj == 0 or int(prog.jobs["cube"].max()) < round_bounds[1]))
        if order is not None:
            full = np.asarray(order, dtype=np.int32)
        else:
            # one round: the static rounds are paired (dispatch_order with G)
            full = dispatch_order(run, self.num_workgroups if one_round else None)
        cubes = prog.jobs["cube"][full] if nj else np.zeros(0, np.int32)
        self.rounds: list[tuple[int, int]] = []
        if one_round:
            self.rounds = [(0, len(full))] if len(full) else []
        else:
            for lo, hi in zip(round_bounds[:-1], round_bounds[1:]):
                a = int(np.searchsorted(cubes, lo, side="left"))
                b = int(np.searchsorted(cubes, hi, side="left"))
                if b > a:
                    self.rounds.append((a, b))
        self.order = torch.from_numpy(np.ascontiguousarray(full, dtype=np.int32)).to(device)
        # One round: the jobs go to the device already in dispatch order and the kernel
        # takes position p as job p (no order lookup before a job starts); job flags are
        # then per position (job_flags() maps them back).
        self.perm = np.ascontiguousarray(full, dtype=np.int64) if one_round and nj else None
        if self.perm is not None:
            permuted = np.ascontiguousarray(self.dp.jobs_host[self.perm])
            self.dp.jobs_host = permuted
            self.dp.jobs = torch.from_numpy(permuted.view(np.uint8).copy()).to(device)
        longest = max((b - a for a, b in self.rounds), default=1)
        self.worklist = torch.empty(max(1, longest), dtype=torch.int32, device=device)
        # Everything a launch resets lives in one buffer, cleared by one fill kernel: per
        # round r (CS int32s) the queue counters (BATCH_COUNTER_INTS) and the compacted
        # length on a line of its own, then the verdicts, then the per-job flags.
        self.cs = N.BATCH_COUNTER_INTS + N.BATCH_COUNTER_STRIDE
        n_cnt = self.cs * max(1, len(self.rounds))
        n_ver = (max(1, prog.num_instances) + 63) // 64 * 64
        n_done = max(1, nj) if track_jobs else 0
        self.resets = torch.zeros(n_cnt + n_ver + n_done, dtype=torch.int32, device=device)
        self.counters = self.resets[:n_cnt]
        self.verdicts = self.resets[n_cnt : n_cnt + max(1, prog.num_instances)]
        self.job_done = self.resets[n_cnt + n_ver :] if track_jobs else None

    def launch(self, mode: str = "GT0", early_exit: bool = True, events=None) -> None:
        """
        Enqueue one full decision pass on the current stream (verdicts reset first).
        ``events`` (start, end) are recorded right around the round kernels.
        """
        L = N.lib()
        self.resets.zero_()  # counters, verdicts and job flags: one fill kernel
        if events is not None:
            events[0].record()
        stream = N.stream_ptr()
        cbase = self.counters.data_ptr()
        for r, (a, b) in enumerate(self.rounds):
            n = b - a
            order_ptr = None if self.perm is not None else self.order.data_ptr() + 4 * a
            num_dev = None
            if r > 0 and early_exit:
                N.check(L.ts_batch_compact(ctypes.c_void_p(order_ptr), n, N.dptr(self.dp.jobs),
                                           N.dptr(self.verdicts), N.dptr(self.worklist),
                                           ctypes.c_void_p(cbase + 4 * (self.cs * r + N.BATCH_COUNTER_INTS)), stream),
                        "ts_batch_compact")
                order_ptr = self.worklist.data_ptr()
                num_dev = cbase + 4 * (self.cs * r + N.BATCH_COUNTER_INTS)
            bt = N.Batch(
                jobs=self.dp.jobs.data_ptr(), order=order_ptr, num_jobs=n,
                steps=self.dp.steps.data_ptr(), leaves=self.dp.leaves.data_ptr(), arenas=self.arenas.data_ptr(),
                arena_stride=self.arena_stride, num_workgroups=self.num_workgroups, mode=N.MODES[mode],
                verdicts=self.verdicts.data_ptr(), counter=cbase + 4 * (self.cs * r), early_exit=int(early_exit),
                job_done=self.job_done.data_ptr() if self.job_done is not None else None,
                num_jobs_dev=num_dev,
                lds_arena=(4 if self.lane_form else 3 if self.packets else 2 if self.one_wave else 1) if self.lds else 0,
            )
            N.check(L.ts_batch_run(ctypes.byref(bt), stream), "ts_batch_run")
        if events is not None:
            events[1].record()

    def job_flags(self) -> np.ndarray:
        """Per job of the program (job order): 1 ran, 2 skipped, 0 not taken, from the last
        launch (needs track_jobs)."""
        if self.job_done is None:
            raise ValueError("job flags need track_jobs=True")
        nj = len(self.prog.jobs)
        f = self.job_done.cpu().numpy()[:nj]
        if self.perm is None:
            return f
        out = np.zeros(nj, dtype=f.dtype)
        out[self.perm] = f
        return out

    def run(self, mode: str = "GT0", early_exit: bool = True) -> np.ndarray:
        torch = _torch()
        self.launch(mode, early_exit)
        torch.cuda.synchronize()
        return self.verdicts.cpu().numpy()[: self.prog.num_instances].astype(np.uint8)


class BatchCounter:
    """
    Exact model counts of all jobs of a program in one persistent launch
    (``ts_batch_count``): uint64 arenas, per-instance sums of result x 2^free_vars,
    overflow flags raised instead of wrapping.
    """

    def __init__(self, prog: Program, *, num_workgroups: int | None = None, device: str = "cuda") -> None:
        torch = _torch()
        L = N.lib()
        self.dp = DeviceProgram(prog, device)
        self.prog = prog
        nj = len(prog.jobs)
        if num_workgroups is None:
            mx = N.i32(0)
            N.check(L.ts_batch_max_workgroups(ctypes.byref(mx)), "ts_batch_max_workgroups")
            num_workgroups = int(mx.value)
        max_arena = int(prog.jobs["arena_elems"].max()) if nj else 16
        self.arena_stride = (8 * max_arena + 255) // 256 * 256
        # at least one workgroup per non-empty queue segment; device arenas within budget
        num_workgroups = min(int(num_workgroups), max(N.BATCH_SEGMENTS, ARENA_BUDGET // self.arena_stride))
        self.num_workgroups = max(1, min(max(num_workgroups, N.BATCH_SEGMENTS), max(1, nj)))
        self.arenas = torch.empty(self.num_workgroups * self.arena_stride, dtype=torch.uint8, device=device)
        self.order = torch.from_numpy(dispatch_order(prog)).to(device)
        n = max(1, prog.num_instances)
        self.counts = torch.zeros(n, dtype=torch.int64, device=device)
        self.overflow = torch.zeros(n, dtype=torch.int32, device=device)
        self.counter = torch.zeros(N.BATCH_COUNTER_INTS, dtype=torch.int32, device=device)
        self.dummy = torch.zeros(n, dtype=torch.int32, device=device)

    def launch(self, events=None) -> None:
        L = N.lib()
        self.counts.zero_()
        self.overflow.zero_()
        self.counter.zero_()
        if events is not None:
            events[0].record()
        b = N.Batch(
            jobs=self.dp.jobs.data_ptr(), order=self.order.data_ptr(), num_jobs=int(self.order.numel()),
            steps=self.dp.steps.data_ptr(), leaves=self.dp.leaves.data_ptr(), arenas=self.arenas.data_ptr(),
            arena_stride=self.arena_stride, num_workgroups=self.num_workgroups, mode=0,
            verdicts=self.dummy.data_ptr(), counter=self.counter.data_ptr(), early_exit=0,
            job_done=None, num_jobs_dev=None,
        )
        N.check(L.ts_batch_count(ctypes.byref(b), N.dptr(self.counts), N.dptr(self.overflow), N.stream_ptr()),
                "ts_batch_count")
        if events is not None:
            events[1].record()

    def run(self) -> list[int]:
        torch = _torch()
        self.launch()
        torch.cuda.synchronize()
        n = self.prog.num_instances
        if n and int(self.overflow[:n].max().item()):
            raise OverflowError("model count exceeds 2^64")
        vals = self.counts.cpu().numpy()[:n].view(np.uint64)
        return [int(v) for v in vals]
