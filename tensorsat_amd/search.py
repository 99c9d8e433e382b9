"""
Batched SAT decisions with the whole cube search on the device (``ts_search_run``).

The batched counterpart of the reference's per-instance decision
``bool(CotengraContraction(FinRel, wiring).contract(diagram))``
(``examples/test_sat_solver.py:5-9``, ``contractions/cotengra.py:100-128``): each
instance's network is sliced on its variable wires by a depth-first branch on the GPU
(one 64-lane wave per job), clause tensors that become all-zero prune a branch and
one-hot ones fix their variable, and every residual of at most ``width`` variables is
contracted into one relation and reduced (``include/tensorsat_amd.h``, device cube
search).  The host only packs the clauses and lists the jobs.

Hard instances are split: a job owns the subtree below a fixed prefix of the first
``split`` branch choices, so 2^split waves search one instance at once, and the jobs are
launched hardest first (an estimate of the search size from the instance's size and
clause ratio); the few instances with the largest estimates get more split levels, since
the longest search sets the batch's time.  With early exit, the jobs of an instance stop
once one of them finds a satisfiable leaf.  Verdicts do not depend on the split or the
order.  :func:`decide_stream` overlaps the host work of the next batches with the search
of the current one.

:class:`SearchSolver` decides instances the kernel's limits admit (at most 128
variables, 768 clauses after normalisation, 3 distinct literals per clause); the others
raise, and :func:`supported` tells them apart beforehand.
"""

from __future__ import annotations

import ctypes
import itertools
import math
import operator
import time

import numpy as np

from . import _native as N

MAX_VARS = 128
SLOTS = 12
MAX_CLAUSES = 64 * SLOTS
MAX_WIDTH = 6
MAX_SPLIT = 31
QUEUE_STATE = 256  # int32 words of a queue launch's state (TS_SEARCH_QUEUE_STATE)
SAT = 1
INCOMPLETE = 2

INST_DTYPE = np.dtype([("clause_begin", "<i4"), ("num_clauses", "<i4"), ("num_vars", "<i4"), ("reserved", "<i4")])
JOB_DTYPE = np.dtype([("instance", "<i4"), ("depth", "<i4"), ("split", "<i4"), ("own_from", "<i4"),
                      ("prefix", "<u4"), ("reserved", "<u4", (3,)), ("assigned", "<u8", (2,)), ("value", "<u8", (2,))])
STAT_DTYPE = np.dtype([("status", "<i4"), ("max_depth", "<i4"), ("nodes", "<i8"), ("leaves", "<i8"),
                       ("conflicts", "<i8"), ("digest", "<u8")])
assert INST_DTYPE.itemsize == 16 and JOB_DTYPE.itemsize == 64 and STAT_DTYPE.itemsize == 40


class _Search(ctypes.Structure):
    _fields_ = [
        ("clauses", ctypes.c_void_p), ("instances", ctypes.c_void_p), ("num_instances", ctypes.c_int32),
        ("jobs", ctypes.c_void_p), ("num_jobs", ctypes.c_int32), ("verdicts", ctypes.c_void_p),
        ("stats", ctypes.c_void_p), ("width", ctypes.c_int32), ("exhaustive", ctypes.c_int32),
        ("early_exit", ctypes.c_int32), ("spill_capacity", ctypes.c_int32), ("max_nodes", ctypes.c_int64),
        ("spill", ctypes.c_void_p), ("spill_count", ctypes.c_void_p), ("max_clauses", ctypes.c_int32),
        ("order", ctypes.c_int32), ("witness", ctypes.c_void_p), ("counts", ctypes.c_void_p),
        ("count_overflow", ctypes.c_void_p),
    ]


def supported(cnf) -> bool:
    """Whether the device search takes this instance (before normalisation limits on
    the clause count, which :func:`pack_batch` checks exactly)."""
    lens, _ = cnf.literal_arrays()
    return cnf.num_vars <= MAX_VARS and (lens.size == 0 or int(lens.max()) <= 3)


def pack_batch(cnfs, alloc=None) -> tuple[np.ndarray, np.ndarray]:
    """(packed clauses, per-instance records) of a batch, by the host library's packer
    (``tsc_search_pack``: duplicate literals dropped, tautologies removed, clause order
    kept); raises NotImplementedError for an instance outside the kernel's limits and
    ValueError for a literal out of range.  ``alloc(words, instances)``, if given, returns
    the (uint32 [>= words], INST_DTYPE [instances]) arrays to pack into."""
    from . import program as PG

    L = PG._lib()
    num = len(cnfs)
    try:  # every header already cached: gathered at C speed
        hb = b"".join(map(_HEADER_BYTES, cnfs))
    except (AttributeError, TypeError):
        hb = b"".join(c.pack_header_bytes() for c in cnfs)
    head = np.frombuffer(hb, dtype=np.int64).reshape(num, 4)
    nv = head[:, 0].astype(np.int32)
    ncl = np.ascontiguousarray(head[:, 1])
    lens_p = np.ascontiguousarray(head[:, 2])
    lits_p = np.ascontiguousarray(head[:, 3])
    if alloc is not None:
        words, inst = alloc(max(1, int(ncl.sum())), num)
    else:
        words = np.empty(max(1, int(ncl.sum())), dtype=np.uint32)
        inst = np.zeros(num, dtype=INST_DTYPE)
    total, bad = ctypes.c_int64(0), ctypes.c_int32(-1)
    rc = L.tsc_search_pack(num, nv.ctypes.data, ncl.ctypes.data, lens_p.ctypes.data, lits_p.ctypes.data,
                           words.ctypes.data, inst.ctypes.data if num else None, ctypes.byref(total),
                           ctypes.byref(bad), 0)
    if rc == -3:
        raise NotImplementedError(f"instance {bad.value} is outside the device search's limits ({MAX_VARS} "
                                  f"variables, {MAX_CLAUSES} clauses, 3 distinct literals per clause)")
    if rc != 0:
        raise ValueError(f"instance {bad.value}: literal out of range")
    return words[: total.value], inst


_HEADER_BYTES = operator.attrgetter("_hb")


def _align(n: int) -> int:
    return (n + 255) & ~255


def _decode(v: np.ndarray) -> np.ndarray:
    """Verdict words -> 1 SAT, 0 UNSAT, -1 undecided."""
    v = v.astype(np.int64)
    out = np.where(v & SAT, 1, 0)
    out[((v & INCOMPLETE) != 0) & ((v & SAT) == 0)] = -1
    return out


def log2_search_size(num_vars: np.ndarray, num_clauses: np.ndarray) -> np.ndarray:
    """Estimated log2 of the branch nodes and conflicts one search needs with early exit,
    for random 3-SAT-like instances: linear in n with a slope that grows with the clause
    ratio up to the satisfiability threshold (4.26) and falls past it (least squares on
    the batched workload: slope 0.054 at ratio 4.26, 0.024 at 3.5).  Used only to order
    and split jobs."""
    n = np.asarray(num_vars, dtype=np.float64)
    r = np.asarray(num_clauses, dtype=np.float64) / np.maximum(n, 1.0)
    lo, hi = 0.024, 0.054
    slope = np.where(r <= 4.26, lo + (hi - lo) * np.clip((r - 3.5) / 0.76, 0.0, 1.0), hi * 4.26 / np.maximum(r, 1e-9))
    return 2.5 + slope * n


def plan_jobs(inst: np.ndarray, *, split_target: float | None = 128.0, max_split: int = 8,
              tail_fraction: float = 0.0, tail_boost: int = 0, out: np.ndarray | None = None) -> np.ndarray:
    """One job per subtree: instance i gets 2^split_i jobs (prefixes 0..2^split_i - 1)
    where split_i = ceil(log2(estimated nodes / split_target)) (:func:`log2_search_size`;
    0 with ``split_target`` None), launched in decreasing order of estimated size per job
    (instances of equal size in index order, the jobs of an instance in prefix order).
    The instances whose estimate is at least the k-th largest of the batch, k =
    ceil(``tail_fraction`` * count), get ``tail_boost`` more split levels (capped by
    ``max_split``).  Built by the host library (``tsc_search_plan``), into ``out`` when it
    has room."""
    from . import program as PG

    if split_target is not None and split_target <= 0:
        raise ValueError("split_target must be positive")
    num = len(inst)
    if num == 0:
        return np.zeros(0, dtype=JOB_DTYPE)
    L = PG._lib()
    rec = np.ascontiguousarray(inst)
    target = 0.0 if split_target is None else float(split_target)
    cap = 0 if out is None else len(out)
    count = ctypes.c_int64(0)
    buf = out if out is not None else np.zeros(0, dtype=JOB_DTYPE)
    tf, tb = float(tail_fraction), int(tail_boost)
    rc = L.tsc_search_plan(num, rec.ctypes.data, target, max_split, tf, tb, buf.ctypes.data if cap else None, cap,
                           ctypes.byref(count))
    if rc == -2:  # no room: the count is known now
        buf = np.empty(count.value, dtype=JOB_DTYPE)
        rc = L.tsc_search_plan(num, rec.ctypes.data, target, max_split, tf, tb, buf.ctypes.data, count.value,
                               ctypes.byref(count))
    if rc != 0:
        raise ValueError("tsc_search_plan: invalid arguments")
    return buf[: count.value]


def _var_masks(nv: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Per count of variables: the masks of valid bits in the two assignment words."""
    one, ones = np.uint64(1), np.uint64(0xFFFFFFFFFFFFFFFF)
    lo = np.where(nv >= 64, ones, (one << np.clip(nv, 0, 63).astype(np.uint64)) - one)
    hi = np.where(nv <= 64, np.uint64(0),
                  np.where(nv >= 128, ones, (one << np.clip(nv - 64, 0, 63).astype(np.uint64)) - one))
    return lo.astype(np.uint64), hi.astype(np.uint64)


def _popcount64(x: np.ndarray) -> np.ndarray:
    return np.bitwise_count(np.asarray(x, dtype=np.uint64)).astype(np.int64)


def validate(words: np.ndarray, inst: np.ndarray, jobs: np.ndarray | None, width: int, *,
             literals: bool = True) -> None:
    """Host-side bounds check of everything the kernel indexes, before any upload
    (``literals=False`` skips the per-literal variable check, for clause words made by
    :func:`pack_batch`, which checks them as it packs; ``jobs=None`` skips the job checks,
    for start jobs made by :func:`plan_jobs`, whose planner writes only instances of the
    batch, splits of 0..31, prefixes below 2^split and owners at or above the root)."""
    if not 0 <= width <= MAX_WIDTH:
        raise ValueError(f"width must be in 0..{MAX_WIDTH}")
    if len(inst):
        if np.any(inst["num_clauses"] < 0) or np.any(inst["num_clauses"] > MAX_CLAUSES):
            raise ValueError("clause count outside the kernel's limit")
        if np.any(inst["num_vars"] < 0) or np.any(inst["num_vars"] > MAX_VARS):
            raise ValueError("variable count outside the kernel's limit")
        end = inst["clause_begin"].astype(np.int64) + inst["num_clauses"]
        if np.any(inst["clause_begin"] < 0) or np.any(end > words.size):
            raise ValueError("clause range outside the clause array")
        if words.size and literals:
            length = words >> 24
            if np.any(length > 3):
                raise ValueError("packed clause with more than 3 literals")
            counts = inst["num_clauses"].astype(np.int64)
            owner = np.repeat(np.arange(len(inst)), counts)
            begin = np.repeat(inst["clause_begin"].astype(np.int64), counts)
            first = np.repeat(np.cumsum(counts) - counts, counts)
            w = words[begin + np.arange(owner.size) - first]
            nvars = inst["num_vars"][owner].astype(np.int64)
            for i in range(3):
                v = (w >> (8 * i + 1)) & 0x7F
                used = (w >> 24) > i
                if np.any(used & (v >= nvars)):
                    raise ValueError("literal of a variable beyond the instance's count")
    if jobs is not None and len(jobs):
        if np.any(jobs["instance"] >= len(inst)):
            raise ValueError("job names an instance outside the batch")
        if np.any(jobs["split"] < 0) or np.any(jobs["split"] > MAX_SPLIT):
            raise ValueError(f"split outside 0..{MAX_SPLIT}")
        if np.any(jobs["prefix"].astype(np.int64) >> jobs["split"].astype(np.int64)):
            raise ValueError("prefix has bits beyond its split")
        if np.any(jobs["own_from"] < 0) or np.any(jobs["own_from"] > jobs["depth"] + jobs["split"]):
            raise ValueError("job owns from a depth below its prefix levels")
        if not jobs["depth"].any() and not jobs["assigned"].any() and not jobs["value"].any():
            return  # jobs from the root: nothing more to check
        live = jobs["instance"] >= 0
        nv = inst["num_vars"][np.maximum(jobs["instance"], 0)].astype(np.int64)
        lo_mask, hi_mask = _var_masks(nv)
        asg, val = jobs["assigned"], jobs["value"]
        if np.any(live & (((asg[:, 0] & ~lo_mask) != 0) | ((asg[:, 1] & ~hi_mask) != 0))):
            raise ValueError("job assignment names variables beyond the instance's count")
        if np.any(((val[:, 0] & ~asg[:, 0]) != 0) | ((val[:, 1] & ~asg[:, 1]) != 0)):
            raise ValueError("job value bits outside its assigned variables")
        free = nv - _popcount64(asg[:, 0]) - _popcount64(asg[:, 1])
        # every branch below the start assigns a free variable: the branch stack reaches
        # depth + free - 1 at most, which must stay inside its MAX_VARS frames
        if np.any(live & ((jobs["depth"] < 0) | (jobs["depth"] + free > MAX_VARS))):
            raise ValueError("job start depth leaves no room for its branch stack")


class SearchSolver:
    """Decides a batch of CNF instances with the device cube search.

    ``width``: residuals of at most this many variables are contracted (0..6).
    ``split_target``: estimated branch nodes per job (hard instances are split into 2^k
    jobs from the start; None: one job per instance).  ``budget``: nodes a job searches
    before it hands its remaining subtrees over as new jobs (None: no hand-over): into a
    list run by the next launch, launches repeating until no job is handed over (the
    host reads each round's count); or, with ``queue``, into a device queue that the
    same launch drains (one launch, no host round trip -- for callers that cannot
    synchronise; on the batched workload it measures 10-30 % slower than the rounds,
    the claims on one counter costing more than the round trips).  The verdicts do not
    depend on the split, the budget or the mode.  ``exhaustive``: visit every branch (for
    statistics; no early exit).  ``max_nodes``: with ``budget`` None, the per-job node
    cap; an instance whose job hits it is reported undecided (-1) unless another job
    found it SAT."""

    def __init__(self, *, width: int = 6, split_target: float | None = 128.0, max_split: int = 8,
                 budget: int | None = None, exhaustive: bool = False, max_nodes: int = 1 << 26,
                 stats: bool = False, device: str = "cuda", spill_capacity: int | None = None,
                 max_rounds: int = 100_000, order: str = "score", witness: bool = False, queue: bool = False,
                 workers: int = 0, count: bool = False, tail_fraction: float = 0.01, tail_boost: int = 3):
        """``order``: "score" (the decision default) or "lex" (smallest residual variable,
        false first); ``count``: exact model counts (:meth:`counts`; implies exhaustive);
        ``witness``: keep each satisfiable instance's assignment
        (:meth:`witnesses`).  ``tail_fraction`` / ``tail_boost``: the instances of a batch
        with the largest estimated searches (that fraction of them) get that many more
        split levels (:func:`plan_jobs`; on 4096 random 3-SAT instances the batch time
        fell 0.54 -> 0.40 ms and 0.40 -> 0.39 ms on two batches).  "lex" with one job per instance (``split_target`` None,
        ``budget`` None) makes that assignment the lexicographically smallest."""
        if not 0 <= width <= MAX_WIDTH:
            raise ValueError(f"width must be in 0..{MAX_WIDTH}")
        if budget is not None and budget < 1:
            raise ValueError("budget must be >= 1 (or None)")
        if order not in ("score", "lex"):
            raise ValueError("order must be 'score' or 'lex'")
        self.order = order
        self.want_witness = witness
        self.want_count = bool(count)
        exhaustive = exhaustive or self.want_count
        self.queue = bool(queue) and budget is not None
        self.workers = int(workers)
        self.width = width
        self.split_target = split_target
        self.max_split = max_split
        self.tail_fraction = float(tail_fraction)
        self.tail_boost = int(tail_boost)
        self.budget = budget
        self.exhaustive = exhaustive
        self.max_nodes = int(max_nodes)
        self.want_stats = stats or exhaustive
        self.device = device
        self.spill_capacity = spill_capacity
        self.max_rounds = max_rounds
        self.timing: dict = {}
        self._pin = self._dev_stage = self._staged = None
        self._vpin = self._vdone = None
        self.rounds = 0
        self.round_jobs: list[int] = []

    def prepare(self, cnfs, sync: bool = True) -> None:
        """Pack, plan the jobs and upload (host work, timed in ``self.timing``).  Clauses,
        instance records and jobs are written straight into one pinned staging buffer
        (kept by the solver, grown as needed) and sent with one host-to-device copy.
        ``sync`` False returns once the copy is enqueued on the current stream.  The two
        halves are :meth:`stage` and :meth:`submit`."""
        import torch

        st = self.stage(cnfs)
        t2 = time.perf_counter()
        self.submit(st)
        if sync:
            torch.cuda.synchronize()
            self.timing["upload_s"] = time.perf_counter() - t2

    def stage(self, cnfs) -> dict:
        """Host half of :meth:`prepare`: pack and plan into the pinned staging buffer (after
        this solver's previous staging copy has completed).  It reads and writes only the
        staging buffer, so it may run on another thread while this solver's previous batch
        is searched or its verdicts are read; no GPU work is enqueued."""
        if self._staged is not None:
            self._staged.synchronize()  # the previous batch's copy has read the staging buffer
        t0 = time.perf_counter()
        layout = {}

        def alloc(nwords, ninst):
            layout["words"] = (0, nwords * 4)
            layout["inst"] = (_align(nwords * 4), ninst * INST_DTYPE.itemsize)
            self._stage_pin(layout["inst"][0] + ninst * INST_DTYPE.itemsize)
            h = self._pin.numpy()
            return (h[: nwords * 4].view(np.uint32),
                    h[layout["inst"][0] : layout["inst"][0] + layout["inst"][1]].view(INST_DTYPE))

        words, inst = pack_batch(cnfs, alloc)
        t1 = time.perf_counter()
        jo = _align(layout["inst"][0] + layout["inst"][1])
        room = (self._pin.numel() - jo) // JOB_DTYPE.itemsize
        h = self._pin.numpy()
        jobs = plan_jobs(inst, split_target=self.split_target, max_split=self.max_split,
                         tail_fraction=self.tail_fraction, tail_boost=self.tail_boost,
                         out=h[jo : jo + room * JOB_DTYPE.itemsize].view(JOB_DTYPE) if room > 0 else None)
        if room <= 0 or not np.may_share_memory(jobs, h):  # grown: stage the jobs too
            need = jo + jobs.nbytes
            old = (words.copy(), inst.copy())
            self._stage_pin(need)
            h = self._pin.numpy()
            words = h[: old[0].nbytes].view(np.uint32)
            words[:] = old[0]
            inst = h[layout["inst"][0] : layout["inst"][0] + old[1].nbytes].view(INST_DTYPE)
            inst[:] = old[1]
            staged = h[jo : jo + jobs.nbytes].view(JOB_DTYPE)
            staged[:] = jobs
            jobs = staged
        layout["jobs"] = (jo, jobs.nbytes)
        validate(words, inst, None, self.width, literals=False)  # packer and planner checked the rest
        return {"words": words, "inst": inst, "jobs": jobs, "layout": layout, "pin": self._pin,
                "pack_s": t1 - t0, "plan_s": time.perf_counter() - t1}

    def submit(self, st: dict) -> None:
        """Device half of :meth:`prepare`: the staged records to the device with one copy
        enqueued on the current stream (call from the thread that launches)."""
        t2 = time.perf_counter()
        self.upload(st["words"], st["inst"], st["jobs"], staged=st["layout"], pin=st["pin"])
        self.timing.update(pack_s=st["pack_s"], plan_s=st["plan_s"], upload_s=time.perf_counter() - t2)

    def _stage_pin(self, nbytes: int) -> None:
        """Pinned staging buffer of at least ``nbytes`` (its device twin is sized by
        :meth:`upload`)."""
        import torch

        if self._pin is None or self._pin.numel() < nbytes:
            size = max(int(nbytes * 1.25), 1 << 20)
            self._pin = torch.empty(size, dtype=torch.uint8, pin_memory=True)

    def upload(self, words: np.ndarray, inst: np.ndarray, jobs: np.ndarray, staged: dict | None = None,
               pin=None) -> None:
        """Device copies of the records; ``staged`` (from :meth:`stage`): their byte
        ranges in the pinned staging buffer ``pin``, sent with one copy."""
        import torch

        dev = self.device
        self.num_instances = len(inst)
        self.max_clauses = int(inst["num_clauses"].max()) if len(inst) else 0
        self.host_jobs = jobs
        if staged is not None and len(inst) and len(jobs) and words.size:
            end = staged["jobs"][0] + staged["jobs"][1]
            if self._dev_stage is None or self._dev_stage.numel() < end:
                self._dev_stage = torch.empty(max(int(end * 1.25), 1 << 20), dtype=torch.uint8, device=dev)
            pin = self._pin if pin is None else pin
            self._dev_stage[:end].copy_(pin[:end], non_blocking=True)
            self._staged = torch.cuda.Event()
            self._staged.record()

            def view(key):
                a, n = staged[key]
                return self._dev_stage[a : a + n].view(torch.int32)

            self.d_words, self.d_inst, self.d_jobs = view("words"), view("inst"), view("jobs")
        else:
            self.d_words = torch.from_numpy(words.view(np.int32) if words.size else np.zeros(1, np.int32)).to(dev)
            self.d_inst = torch.from_numpy(inst.view(np.int32).reshape(-1) if len(inst)
                                           else np.zeros(4, np.int32)).to(dev)
            self.d_jobs = torch.from_numpy(jobs.view(np.int32).reshape(-1) if len(jobs)
                                           else np.zeros(16, np.int32)).to(dev)
        self.d_verdicts = torch.zeros(max(1, len(inst)), dtype=torch.int32, device=dev)
        self.d_witness = torch.zeros(2 * max(1, len(inst)), dtype=torch.int64, device=dev) \
            if self.want_witness else None
        self.d_counts = torch.zeros(max(1, len(inst)), dtype=torch.int64, device=dev) if self.want_count else None
        self.d_count_ovf = torch.zeros(1, dtype=torch.int32, device=dev) if self.want_count else None
        self.num_vars = inst["num_vars"].copy()
        self._queue_checked = True
        if self.queue:
            # one queue: the start jobs, then room for the hand-overs
            self.cap = int(self.spill_capacity or max(1 << 18, 8 * len(jobs)))
            if self.cap < len(jobs):
                raise ValueError("spill_capacity below the number of start jobs")
            q = np.zeros(self.cap, dtype=JOB_DTYPE)
            q[: len(jobs)] = jobs
            self.d_jobs = torch.from_numpy(q.view(np.int32).reshape(-1)).to(dev)
            self.d_qstate = torch.zeros(QUEUE_STATE, dtype=torch.int32, device=dev)
            self.d_qready = torch.zeros(self.cap, dtype=torch.int32, device=dev)
        elif self.budget is not None:
            cap = self.spill_capacity or max(1 << 18, 8 * len(jobs))
            self.cap = int(cap)
            self.d_spill = [torch.empty(self.cap * JOB_DTYPE.itemsize // 4, dtype=torch.int32, device=dev)
                            for _ in range(2)]
            self.d_count = torch.zeros(1, dtype=torch.int32, device=dev)
        self._stats_rounds: list = []

    def _args(self, jobs_ptr, num_jobs, stats, spill) -> _Search:
        return _Search(
            N.dptr(self.d_words), N.dptr(self.d_inst), self.num_instances, jobs_ptr, num_jobs,
            N.dptr(self.d_verdicts), N.dptr(stats) if stats is not None else None, self.width,
            int(self.exhaustive), int(not self.exhaustive), self.cap if spill is not None else 0,
            self.budget if self.budget is not None else self.max_nodes,
            N.dptr(spill) if spill is not None else None, N.dptr(self.d_count) if spill is not None else None,
            self.max_clauses, 1 if self.order == "lex" else 0,
            N.dptr(self.d_witness) if self.d_witness is not None else None,
            N.dptr(self.d_counts) if self.d_counts is not None else None,
            N.dptr(self.d_count_ovf) if self.d_counts is not None else None,
        )

    def launch(self, events: list | None = None) -> None:
        """Clear the verdicts (a fill kernel) and run the search on the current stream:
        one launch, or with ``budget`` one launch per round of handed-over jobs (the host
        reads each round's job count).  With ``events`` (a list), a pair of timing events
        recorded on that stream around each search launch is appended to it."""
        import torch

        L = N.lib()
        self.d_verdicts.zero_()
        if self.d_witness is not None:
            self.d_witness.zero_()
        if self.d_counts is not None:
            self.d_counts.zero_()
            self.d_count_ovf.zero_()
        self._stats_rounds = []
        if self.queue:
            self._launch_queue(events)
            return
        jobs_ptr, n = N.dptr(self.d_jobs), len(self.host_jobs)
        jobs_tensor = self.d_jobs
        self.round_jobs = []
        rounds = 0
        while n > 0:
            if rounds >= self.max_rounds:
                raise RuntimeError(f"search did not finish in {self.max_rounds} rounds")
            stats = torch.zeros(n * STAT_DTYPE.itemsize // 4, dtype=torch.int32, device=self.device) \
                if self.want_stats else None
            spill = None
            if self.budget is not None:
                spill = self.d_spill[rounds % 2]
                self.d_count.zero_()
            args = self._args(jobs_ptr, n, stats, spill)
            if events is not None:
                pair = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                pair[0].record()
            N.check(L.ts_search_run(ctypes.byref(args), N.stream_ptr()), "ts_search_run")
            if events is not None:
                pair[1].record()
                events.append(pair)
            if stats is not None:
                self._stats_rounds.append((jobs_tensor[: n * JOB_DTYPE.itemsize // 4].clone(), stats))
            self.round_jobs.append(n)
            rounds += 1
            if spill is None:
                break
            n = min(int(self.d_count.item()), self.cap)
            jobs_ptr, jobs_tensor = N.dptr(spill), spill
        self.rounds = rounds

    def _launch_queue(self, events) -> None:
        import torch

        n = len(self.host_jobs)
        stats = torch.zeros(self.cap * STAT_DTYPE.itemsize // 4, dtype=torch.int32, device=self.device) \
            if self.want_stats else None
        args = self._args(N.dptr(self.d_jobs), n, stats, None)
        if events is not None:
            pair = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            pair[0].record()
        N.check(N.lib().ts_search_run_queue(ctypes.byref(args), N.dptr(self.d_jobs), self.cap, N.dptr(self.d_qstate),
                                            N.dptr(self.d_qready), self.workers, N.stream_ptr()), "ts_search_run_queue")
        if events is not None:
            pair[1].record()
            events.append(pair)
        self._queue_stats = stats
        self._queue_checked = False
        self.round_jobs = [n]
        self.rounds = 1

    def queue_records(self) -> int:
        """Queue mode: records the last launch reserved (start jobs and hand-overs; those
        beyond the capacity were searched in place).  Raises RuntimeError if a workgroup
        gave up waiting for a record."""
        st = self.d_qstate.cpu().numpy()
        if st[192]:
            raise RuntimeError("device search queue: a workgroup gave up waiting for a record")
        self._queue_checked = True
        return int(st[0])

    def verdicts(self) -> np.ndarray:
        """Per instance: 1 SAT, 0 UNSAT, -1 undecided (a job ran out of its node budget)."""
        if self.queue and not self._queue_checked:
            self.queue_records()
        return _decode(self.d_verdicts.cpu().numpy()[: self.num_instances])

    def fetch_verdicts(self) -> None:
        """Enqueue the copy of the verdicts into pinned host memory on the current stream
        (after the search launched last); :meth:`fetched_verdicts` waits for it."""
        import torch

        n = max(1, self.num_instances)
        if self._vpin is None or self._vpin.numel() < n:
            self._vpin = torch.empty(max(n, 4096), dtype=torch.int32, pin_memory=True)
        self._vpin[:n].copy_(self.d_verdicts[:n], non_blocking=True)
        self._vdone = torch.cuda.Event()
        self._vdone.record()

    def fetched_verdicts(self) -> np.ndarray:
        """The verdicts copied by the last :meth:`fetch_verdicts`, once that copy is done."""
        if self._vdone is None:
            raise RuntimeError("fetch_verdicts() was not called")
        self._vdone.synchronize()
        return _decode(self._vpin.numpy()[: self.num_instances])

    def counts(self) -> list[int]:
        """Per instance: the exact number of satisfying assignments over its variables
        (construct with ``count=True``).  Raises OverflowError when a count does not fit in
        64 bits and RuntimeError when a job ran out of its node budget."""
        if self.d_counts is None:
            raise ValueError("construct the solver with count=True")
        if int(self.d_count_ovf.item()):
            raise OverflowError("a model count does not fit in 64 bits")
        v = self.d_verdicts.cpu().numpy()[: self.num_instances]
        if np.any(v & INCOMPLETE):
            raise RuntimeError("a search job ran out of its node budget: counts are incomplete")
        c = self.d_counts.cpu().numpy().view(np.uint64)[: self.num_instances]
        return [int(x) for x in c]

    def witnesses(self) -> list:
        """Per instance: the assignment of the satisfiable leaf found ("0"/"1" per
        variable, in index order), or None for an instance not found satisfiable."""
        if self.d_witness is None:
            raise ValueError("construct the solver with witness=True")
        v = self.verdicts()
        w = self.d_witness.cpu().numpy().view(np.uint64).reshape(-1, 2)
        out = []
        for i in range(self.num_instances):
            if v[i] != 1:
                out.append(None)
                continue
            n = int(self.num_vars[i])
            bits = int(w[i, 0]) | (int(w[i, 1]) << 64)
            out.append("".join("1" if (bits >> j) & 1 else "0" for j in range(n)))
        return out

    def job_stats(self) -> tuple[np.ndarray, np.ndarray]:
        """(jobs, statistics) of every job of every round of the last launch."""
        if not self.want_stats:
            raise ValueError("construct the solver with stats=True")
        if self.queue:
            n = min(self.queue_records(), self.cap)
            jobs = self.d_jobs.cpu().numpy().view(JOB_DTYPE)[:n]
            return jobs, self._queue_stats.cpu().numpy().view(STAT_DTYPE)[:n]
        jobs = [j.cpu().numpy().view(JOB_DTYPE) for j, _ in self._stats_rounds]
        st = [s.cpu().numpy().view(STAT_DTYPE) for _, s in self._stats_rounds]
        return np.concatenate(jobs), np.concatenate(st)

    def instance_stats(self) -> np.ndarray:
        """Per-instance sums of the jobs' statistics (nodes, leaves, conflicts, digest
        mod 2^64; max_depth the maximum; status the maximum of complete / stopped /
        exhausted, hand-overs counted as complete)."""
        jobs, js = self.job_stats()
        live = jobs["instance"] >= 0
        jobs, js = jobs[live], js[live]
        out = np.zeros(self.num_instances, dtype=STAT_DTYPE)
        inst = jobs["instance"]
        for f in ("nodes", "leaves", "conflicts"):
            np.add.at(out[f], inst, js[f])
        np.add.at(out["digest"], inst, js["digest"])
        np.maximum.at(out["max_depth"], inst, js["max_depth"])
        np.maximum.at(out["status"], inst, np.where(js["status"] == 3, 0, js["status"]))
        return out

    def run(self, cnfs) -> np.ndarray:
        import torch

        self.prepare(cnfs)
        self.launch()
        torch.cuda.synchronize()
        return self.verdicts()


def decide_stream(batches, *, ahead: int = 3, **solver_kw):
    """Verdicts of a stream of batches (an iterable of CNF lists), one array per batch, in
    order.  ``ahead`` batches are packed and planned on worker threads (:meth:`SearchSolver.stage`;
    the native packer and job planner release the interpreter lock) while the calling thread
    uploads, launches and reads back (:meth:`SearchSolver.submit`, ``launch``,
    ``fetch_verdicts``): each batch's copy, search and verdict copy are enqueued on the
    current stream without waiting, and its verdicts are read once the next batch is
    enqueued behind it.  ``ahead=0`` stages on the calling thread (4096-instance batches
    of random 3-SAT on one MI355X, 21 780 jobs each: 1.14 ms per batch; ahead 1: 1.49 ms,
    the one worker stages while the caller waits; 2: 0.99 ms; 3: 0.94 ms, the Python glue
    of staging and submission holding the interpreter lock is what remains).  The solvers
    form a ring
    of ``ahead + 2``: the one that stages batch k + ahead searched batch k - 2, whose
    verdicts were already returned.  One launch per batch, so ``budget`` must be None
    (hand-over rounds read a job count on the host) and ``queue`` False."""
    import collections
    from concurrent.futures import ThreadPoolExecutor

    if solver_kw.get("budget") is not None or solver_kw.get("queue"):
        raise ValueError("decide_stream needs budget=None and queue=False")
    ahead = max(0, int(ahead))
    solvers = [SearchSolver(**solver_kw) for _ in range(ahead + 2)]
    it = iter(batches)
    staged = collections.deque()
    k = 0
    with ThreadPoolExecutor(max_workers=max(1, ahead)) as ex:

        def schedule() -> None:
            nonlocal k
            for cnfs in it:
                s = solvers[k % len(solvers)]
                k += 1
                staged.append((s, ex.submit(s.stage, cnfs) if ahead else s.stage(cnfs)))
                return

        for _ in range(max(1, ahead)):
            schedule()
        pending = None
        while staged:
            s, st = staged.popleft()
            s.submit(st.result() if ahead else st)
            s.launch()
            s.fetch_verdicts()
            schedule()
            if pending is not None:
                yield pending.fetched_verdicts()
            pending = s
        if pending is not None:
            yield pending.fetched_verdicts()


def autotune(cnfs, *, widths=(4, 6), split_targets=(None, 32.0, 8.0), budgets=(None, 64, 16),
             max_split: int = 8, reps: int = 3, tail_fraction: float = 0.01,
             tail_boosts=(3, 0)) -> tuple[dict, dict]:
    """Pick the residual width, the split target, the tail boost (extra split levels for
    the ``tail_fraction`` of instances with the largest estimated searches) and the
    hand-over budget for a batch: every candidate decides the batch ``reps`` times on the
    GPU (device-resident inputs, wall time of each complete decision, hand-over rounds
    included); the fastest mean wins.  Verdicts of every candidate must agree (they are
    the same decisions), else ``RuntimeError``.  Returns (choice, table)."""
    import torch

    words, inst = pack_batch(cnfs)
    table = {}
    ref = None
    for w, tgt, tb in itertools.product(widths, split_targets, tail_boosts):
        if tgt is None and tb != tail_boosts[0]:
            continue  # no split: the boost has nothing to add to
        jobs = plan_jobs(inst, split_target=tgt, max_split=max_split, tail_fraction=tail_fraction,
                         tail_boost=tb)
        validate(words, inst, jobs, w)
        for bud in budgets:
            s = SearchSolver(width=w, split_target=tgt, max_split=max_split, budget=bud,
                             tail_fraction=tail_fraction, tail_boost=tb)
            s.upload(words, inst, jobs)
            s.launch()
            torch.cuda.synchronize()
            v = s.verdicts()
            if ref is None:
                ref = v
            elif not np.array_equal(v, ref):
                raise RuntimeError(f"verdicts differ between search configurations ({w}, {tgt}, {tb}, {bud})")
            ts = []
            for _ in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                s.launch()
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            table[(w, tgt, tb, bud)] = {"width": w, "split_target": tgt, "tail_boost": tb, "budget": bud,
                                        "jobs": int(len(jobs)), "rounds": s.rounds,
                                        "max_split": int(jobs["split"].max()) if len(jobs) else 0,
                                        "ms": float(np.mean(ts)) * 1000.0}
    best = min(table.values(), key=lambda r: r["ms"])
    return {"width": best["width"], "split_target": best["split_target"], "tail_boost": best["tail_boost"],
            "budget": best["budget"]}, table
