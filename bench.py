#!/usr/bin/env python3
"""
bench.py — batched random 3-SAT decisions on one MI355X.

Default workload: 4096 random 3-SAT instances (instance i from default_rng(10000 + i),
n ~ U{20..80}, clause ratio 4.26 for even i and 3.5 for odd i) decided by the device cube
search (``tensorsat_amd.search``): the packed clauses are resident in HBM and one timed
"step" is the whole decision of the batch on the GPU (verdict reset + one launch of the
search kernel: conditioned slicing, unit propagation and the contraction of every
residual of at most W variables), with no host work.  The end-to-end rate from the list
of CNF instances (packing, job planning, upload, launch, verdicts back) is reported beside
it, with the search width and split chosen by the autotuner.  Verdicts are checked
against the oracle's restatement of the search (``oracle/search.c``), the
reference-equivalent path (the host cube planner and the oracle's numpy contraction of
every cube, on all host cores: the timed ``cpu_baseline``) and an independent DPLL.  The
same line nests the single n=120 decision and the n=100 #SAT count, each with its own
roofline and CPU baseline.

``--workload batched-host`` runs the earlier pipeline (host cube planning, persistent
contraction kernel); ``single``, ``count``, ``learned``, ``dense``, ``rewire`` and
``copy`` run one workload on its own.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

METRIC = "instances/s batched 3-SAT (bit-exact) ; plan+contract s, learned vs plain"
HBM_PEAK_GBS = 8000.0

# HBM bytes per timed step measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled
# by the copy calibration: FETCH x 2.0, WRITE x 1.0) for the configurations profiled;
# other configurations report null.
MEASURED_TRAFFIC = {
    ("batched", 6, 4096, 4): (44.4e6, "profiles/k_batch_wave_batched3sat_w6.md"),
    ("search", 6, 4096): (26.4e6, "profiles/k_search_batched3sat.md"),
    ("dense", 13, 13, 13, 6): (1.120e9, "profiles/k_step_mfma_dense_composition_13_13_13.md"),
    ("dense", 13, 13, 13, 4): (1.468e9, "profiles/k_step_mfma_dense_composition_13_13_13.md"),
}


def batch_kernel_name(ex) -> str:
    """The persistent kernel a BatchExecutor launches."""
    if not ex.lds:
        return "k_batch"
    if ex.one_wave:
        if ex.lane_form:
            return "k_batch_wave<2> (lane form)"
        return "k_batch_wave<1> (packet form)" if ex.packets else "k_batch_wave<0>"
    return "k_batch<true>"


def job_bytes(prog) -> np.ndarray:
    """Algorithmic bytes of each job: leaves once, then every step reads its operands
    once and writes its output once (1 byte per element)."""
    steps = prog.steps
    one = np.int64(1)
    if len(steps) == 0:
        return prog.jobs["num_leaf_bytes"].astype(np.int64)
    norhs = (steps["flags"] & 2) != 0
    per = (one << steps["lhs_rank"].astype(np.int64)) + (one << steps["out_rank"].astype(np.int64))
    per = per + np.where(norhs, 0, one << steps["rhs_rank"].astype(np.int64))
    csum = np.concatenate([[0], np.cumsum(per)])
    b = prog.jobs["step_begin"].astype(np.int64)
    e = b + prog.jobs["num_steps"].astype(np.int64)
    return csum[e] - csum[b] + prog.jobs["num_leaf_bytes"].astype(np.int64)


def cpu_baseline(solver, n_inst: int, mode: str, seconds: float):
    """Oracle step loop on the same cubes (both passes, cube order, stop at the first
    satisfiable cube), instances in order until the time budget is used; one thread."""
    from oracle import executor as OX

    per_inst: list[list[tuple[int, int]]] = [[] for _ in range(n_inst)]
    for p_idx, (prog, sub) in enumerate(zip(solver.progs, solver.subsets)):
        order = np.lexsort((prog.jobs["cube"], prog.jobs["instance"]))
        for j in order:
            per_inst[int(sub[int(prog.jobs["instance"][j])])].append((p_idx, int(j)))
    verdicts = {}
    t0 = time.perf_counter()
    for i in range(n_inst):
        v = 0
        for p_idx, j in per_inst[i]:
            res, _ = OX.run_job(solver.progs[p_idx], j, mode)
            if res.reshape(-1)[0]:
                v = 1
                break
        verdicts[i] = v
        if time.perf_counter() - t0 >= seconds:
            break
    return verdicts, time.perf_counter() - t0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--width", default="auto",
                    help="cube width W (max residual width), or 'auto': autotuned on a 512-instance sample")
    ap.add_argument("--first-pass", type=int, default=4, help="cubes per instance in the first pass")
    ap.add_argument("--mode", default="GT0", choices=["GT0", "MOD256"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workgroups", type=int, default=0)
    ap.add_argument("--threads", type=int, default=0, help="host planning threads")
    ap.add_argument("--rounds", default="", help="dispatch round boundaries by cube index, e.g. 0,4,64 "
                    "(default: the executor's)")
    ap.add_argument("--kernel", default="auto", choices=["auto", "lds", "global"],
                    help="batched kernel: LDS arenas, device-memory arenas, or auto (LDS when the arenas fit)")
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "batched-host", "dense", "learned", "copy", "single", "count", "rewire", "plain", "gnn"])
    ap.add_argument("--split", default="auto", help="batched: split target in branch nodes per job, 'none', "
                    "or 'auto' (autotuned with the width)")
    ap.add_argument("--no-nested", action="store_true", help="batched: skip the nested single / count lines")
    ap.add_argument("--budget-nodes", default="auto", help="batched: nodes a search job runs before it hands "
                    "its remaining subtrees over, 'none', or 'auto'")
    ap.add_argument("--seeds", type=int, default=3, help="single / count workloads: instances (seeds 0..)")
    ap.add_argument("--gang", type=int, default=None,
                    help="single / plain: largest gang of cubes (slices) advanced together (default 8 for the "
                         "n=120 cubes, 32 for the n=40 slices; 0: one at a time on --streams streams)")
    ap.add_argument("--streams", type=int, default=4, help="single / plain with --gang 0: streams")
    ap.add_argument("--min-sum-bits", type=int, default=6,
                    help="dense steps with at least this many summed bits take the int8 MFMA route (0..27)")
    ap.add_argument("--learned-instances", type=int, default=8)
    ap.add_argument("--learned-width", type=int, default=22, help="learned workload: cube width W")
    ap.add_argument("--learned-k", type=int, default=0, help="learned workload: keep only clauses of this k (0: all)")
    ap.add_argument("--budget", type=float, default=8.0, help="learned workload: seconds per instance and method")
    ap.add_argument("--dense-bits", type=int, nargs=3, default=[13, 13, 13], metavar=("M", "N", "K"),
                    help="dense workload: log2 rows, columns and contracted extent")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(instances and cube slices would shard across GPUs with one final OR).", file=sys.stderr)
        return 2

    if args.workload == "dense":
        return bench_dense(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.min_sum_bits != 6:
        import torch

        from tensorsat_amd import _native as N

        torch.zeros(1, device="cuda")  # the device context first
        N.check(N.lib().ts_dense_config(1, args.min_sum_bits), "ts_dense_config")
    if args.workload == "plain":
        return bench_plain(args)
    if args.workload == "single":
        return bench_single(args)
    if args.workload == "gnn":
        def log(msg):
            print(f"[bench] {msg}", file=sys.stderr, flush=True)

        line = run_gnn(args, log)
        line.update({"n_gpus": 1, "warmup": 2, "scaling": "weak", "vs_baseline": None, "data": "synthetic"})
        print(json.dumps(line), flush=True)
        return 0
    if args.workload == "rewire":
        return bench_rewire(args)
    if args.workload == "count":
        return bench_count(args)
    if args.workload == "batched":
        return bench_search(args)
    return bench_batched_host(args)


def host_cores() -> int:
    """Host threads this run may use: the CPU share of one GPU on a shared box (the
    OMP_NUM_THREADS the box sets), else the affinity mask."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(avail, int(env))) if env and env.isdigit() else max(1, avail)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "host CPU"


# ---- process pool over the oracle -----------------------------------------------------
# Forked workers inherit the program being checked (no copy through a pipe); they only
# run numpy and leave through os._exit, never touching the GPU state they inherit.
_POOL_PROG = None


def _pool_decide(job_lists, mode):
    """Verdicts of instances, each given as its job indices in cube order (early exit)."""
    from oracle import executor as OX

    out = []
    for jobs in job_lists:
        v = 0
        for j in jobs:
            res, _ = OX.run_job(_POOL_PROG, j, mode)
            if res.reshape(-1)[0]:
                v = 1
                break
        out.append(v)
    return out


def _pool_job(j, mode):
    from oracle import executor as OX

    res, _ = OX.run_job(_POOL_PROG, j, mode)
    return int(res.reshape(-1)[0])


def _pool(prog, cores):
    import multiprocessing as mp
    import types
    from concurrent.futures import ProcessPoolExecutor

    global _POOL_PROG
    _POOL_PROG = types.SimpleNamespace(jobs=prog.jobs, steps=prog.steps, leaves=prog.leaves,
                                       num_instances=prog.num_instances)
    return ProcessPoolExecutor(max_workers=cores, mp_context=mp.get_context("fork"))


def reference_path_cpu(cnfs, cores: int, width: int = 6, first_pass: int = 4, mode: str = "GT0"):
    """The reference-equivalent CPU path on all host cores: the host cube planner (C++,
    ``cores`` threads) slices every instance into cubes of width <= W, and the oracle's
    numpy step loop (``FinRel._contract2`` semantics, one process per core) contracts
    them in cube order with early exit -- first the first ``first_pass`` cubes of each
    instance, then the remaining cubes of the instances left undecided.  Returns
    (verdicts, seconds, plan seconds)."""
    from tensorsat_amd import program as PG

    t0 = time.perf_counter()
    verdicts = np.zeros(len(cnfs), dtype=np.int64)
    plan_s = 0.0
    todo = np.arange(len(cnfs))
    begin = 0
    for cube_cap in (first_pass, 0):
        if todo.size == 0:
            break
        tp = time.perf_counter()
        prog = PG.sat_batch([cnfs[i] for i in todo], width=width, max_cubes=cube_cap, cube_begin=begin,
                            threads=cores, plan_trials=1)
        plan_s += time.perf_counter() - tp
        order = np.lexsort((prog.jobs["cube"], prog.jobs["instance"]))
        per = [[] for _ in range(len(todo))]
        for j in order:
            per[int(prog.jobs["instance"][j])].append(int(j))
        chunks = [list(range(k, len(todo), cores * 4)) for k in range(cores * 4)]
        with _pool(prog, cores) as ex:
            futs = [ex.submit(_pool_decide, [per[i] for i in ch], mode) for ch in chunks]
            for ch, f in zip(chunks, futs):
                for i, v in zip(ch, f.result()):
                    verdicts[todo[i]] |= v
        truncated = prog.stats[:, 3] == 1
        todo = todo[(verdicts[todo] == 0) & truncated]
        begin = first_pass
    return verdicts, time.perf_counter() - t0, plan_s


def bench_search(args) -> int:
    """The default workload: batched decisions by the device cube search (module
    docstring)."""
    import torch

    from tensorsat_amd import search as SE
    from tensorsat_amd.workloads import batched_3sat

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    cnfs = batched_3sat(args.instances)
    log(f"generated {len(cnfs)} instances in {time.perf_counter() - t0:.1f}s")
    tune = None
    if args.width == "auto" or args.split == "auto":
        t0 = time.perf_counter()
        widths = (4, 6) if args.width == "auto" else (int(args.width),)
        splits = (None, 32.0, 8.0, 2.0) if args.split == "auto" else \
            ((None,) if args.split == "none" else (float(args.split),))
        budgets = (None, 128, 32, 8) if args.budget_nodes == "auto" else \
            ((None,) if args.budget_nodes == "none" else (int(args.budget_nodes),))
        choice, table = SE.autotune(cnfs, widths=widths, split_targets=splits, budgets=budgets)
        width, split, budget = choice["width"], choice["split_target"], choice["budget"]
        boost = choice["tail_boost"]
        tune = {"chosen": choice, "seconds": time.perf_counter() - t0, "candidates": list(table.values())}
        log(f"autotuned width {width}, split target {split}, tail boost {boost}, budget {budget} in "
            f"{tune['seconds']:.1f}s: " + ", ".join(f"W{r['width']}/{r['split_target']}/+{r['tail_boost']}/{r['budget']}: "
                                                    f"{r['ms']:.2f}ms" for r in table.values()))
    else:
        width = int(args.width)
        split = None if args.split == "none" else float(args.split)
        budget = None if args.budget_nodes in ("none", "auto") else int(args.budget_nodes)
        boost = 3
    solver = SE.SearchSolver(width=width, split_target=split, budget=budget, tail_boost=boost)
    # End to end from the CNF list: pack, plan the jobs, upload, launch, verdicts back.
    e2e = []
    for _ in range(5):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver.prepare(cnfs)
        solver.launch()
        v_e2e = solver.verdicts()
        e2e.append(time.perf_counter() - t0)
    e2e_s = float(np.median(e2e))
    host = dict(solver.timing)
    for _ in range(args.warmup):
        solver.launch()
    torch.cuda.synchronize()
    # Timed steps: the whole decision of the batch from device-resident clauses.
    t_start = time.perf_counter()
    for _ in range(args.steps):
        solver.launch()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1000.0 / args.steps
    # Kernel time: events on the launch stream around each k_search launch (a loop of its
    # own): per step the sum over its rounds, per launch the mean over every launch.
    per_step = []
    for _ in range(args.steps):
        ev = []
        solver.launch(events=ev)
        per_step.append(ev)
    torch.cuda.synchronize()
    launch_ms = [[a.elapsed_time(b) for a, b in ev] for ev in per_step]
    kernel_ms = float(np.mean([sum(x) for x in launch_ms]))
    launch_avg_us = 1000.0 * float(np.mean([t for x in launch_ms for t in x]))
    round_us = [round(1000.0 * float(np.mean([x[r] for x in launch_ms if r < len(x)])), 1)
                for r in range(max(len(x) for x in launch_ms))]
    verdicts = solver.verdicts()
    value = len(cnfs) / (ms_per_step / 1000.0)
    jobs = solver.host_jobs
    # Streamed end to end: the batch submitted 16 times in a row from the CNF list; two
    # solvers alternate so the host packs, plans and stages submission k+1 while the GPU
    # searches submission k (one launch per batch: no hand-over budget).
    stream = None
    if budget is None:
        reps = 16
        list(SE.decide_stream([cnfs] * 2, width=width, split_target=split, tail_boost=boost))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs = list(SE.decide_stream([cnfs] * reps, width=width, split_target=split, tail_boost=boost))
        torch.cuda.synchronize()
        per = (time.perf_counter() - t0) / reps
        stream = {"s_per_batch": per, "instances_per_s": len(cnfs) / per, "batches": reps,
                  "mismatches": int(sum(int((v != verdicts).sum()) for v in outs))}
        log(f"streamed end to end: {per * 1000:.2f} ms per batch = {len(cnfs) / per / 1e6:.2f} M inst/s "
            f"({reps} submissions, {stream['mismatches']} verdict mismatches)")
    # Critical path: the statistics of one more launch (per job branch nodes).
    st = SE.SearchSolver(width=width, split_target=split, budget=budget, stats=True, tail_boost=boost)
    st.prepare(cnfs)
    st.launch()
    all_jobs, js = st.job_stats()
    nodes_total = int(js["nodes"].sum())
    busiest = int(np.argmax(js["nodes"] + js["conflicts"]))
    log(f"search: width {width}, split target {split}, budget {budget}: {len(jobs)} jobs (max split "
        f"{int(jobs['split'].max())}), {solver.rounds} rounds ({sum(solver.round_jobs)} jobs), "
        f"{ms_per_step:.3f} ms/step (kernel {kernel_ms:.3f} ms) = {value / 1e6:.2f} M inst/s; end to end "
        f"{e2e_s * 1000:.1f} ms = {len(cnfs) / e2e_s / 1e3:.1f} k inst/s; {nodes_total} branch nodes, "
        f"busiest job {int(js['nodes'][busiest])} nodes")
    nclauses = int(solver.d_words.numel())
    # algorithmic bytes per step: the packed clauses (4 B each) and instance records (16 B)
    # once, every job record (64 B) read once in its round and, for handed-over jobs,
    # written once in the round before, the verdicts (4 B) once.  Over the summed k_search
    # durations of the step this is the mean bytes per launch over the mean launch time.
    alg_bytes = (4.0 * nclauses + 16.0 * len(cnfs) + 64.0 * sum(solver.round_jobs)
                 + 64.0 * sum(solver.round_jobs[1:]) + 4.0 * len(cnfs))
    achieved = alg_bytes / (kernel_ms / 1000.0) / 1e9

    cores = host_cores()
    cpu = None
    parity = {"checked": len(cnfs), "oracle_search_mismatches": None, "reference_path_mismatches": None,
              "dpll_mismatches": None}
    extra = {}
    if not args.no_cpu:
        from oracle import dpll
        from oracle import search as OS

        ref_v, ref_s, ref_plan = reference_path_cpu(cnfs, cores)
        parity["reference_path_mismatches"] = int((ref_v != verdicts).sum())
        cpu = {"value": len(cnfs) / ref_s, "unit": "instances/s", "cores": cores, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"all {len(cnfs)} instances: host cube planner (width 6, {cores} threads, "
                         f"{ref_plan:.2f}s) + the oracle's numpy contraction of the cubes with early exit "
                         f"({cores} processes), {ref_s:.2f}s wall"}
        log(f"cpu reference path: {cpu['value']:.0f} inst/s on {cores} cores ({cpu_model()}); "
            f"mismatches {parity['reference_path_mismatches']}")
        t0 = time.perf_counter()
        osr = OS.search(cnfs, width=width, threads=cores)
        os_s = time.perf_counter() - t0
        parity["oracle_search_mismatches"] = int((osr["verdict"] != verdicts).sum())
        extra["cpu_search_restatement"] = {"value": len(cnfs) / os_s, "unit": "instances/s", "cores": cores,
                                           "sample": f"all {len(cnfs)} instances, oracle/search.c (the same "
                                                     f"search on the CPU, one instance per thread), {os_s:.2f}s"}
        log(f"cpu search restatement: {len(cnfs) / os_s:.0f} inst/s; mismatches {parity['oracle_search_mismatches']}")
        dp = np.array([dpll.solve(c.num_vars, c.clauses) for c in cnfs], dtype=np.int64)
        parity["dpll_mismatches"] = int((dp != verdicts).sum())
        log(f"independent DPLL cross-check: {parity['dpll_mismatches']} mismatches over {len(cnfs)}")
    parity["end_to_end_mismatches"] = int((v_e2e != verdicts).sum())
    parity["streamed_mismatches"] = stream["mismatches"] if stream else None
    parity["undecided"] = int((verdicts < 0).sum())

    nested = {}
    if not args.no_nested:
        nested["plain_n40"] = run_plain(args, seeds=range(3), log=log)
        nested["single_n120"] = run_single(args, seeds=range(3), log=log)
        nested["count_n100"] = run_count(args, seeds=range(1), log=log)
        nested["gnn_train"] = run_gnn(args, log=log)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": "batched random 3-SAT decisions: 4096 instances, n~U{20..80}, ratio 4.26/3.5 "
                        "(default_rng(10000+i)); device cube search (conditioned slicing, unit propagation, "
                        "residuals of <= W variables contracted and reduced) from device-resident packed clauses",
            "instances": len(cnfs),
            "width": width,
            "split_target": split,
            "tail": {"fraction": solver.tail_fraction, "boost": solver.tail_boost},
            "budget": budget,
            "rounds": solver.rounds,
            "jobs_per_round": solver.round_jobs,
            "jobs": int(len(jobs)),
            "max_split": int(jobs["split"].max()),
            "autotune": tune,
            "kernel": "k_search",
            "kernel_ms": kernel_ms,
            "k_search_launch_avg_us": launch_avg_us,
            "k_search_round_us": round_us,
            "launch": "per step: verdict reset (fill kernel) + one k_search launch of one 64-lane wave per job, "
                      "repeated for each round of handed-over subtrees (the host reads each round's job count)",
            "sat": int((verdicts == 1).sum()),
            "end_to_end_s": e2e_s,
            "end_to_end_instances_per_s": len(cnfs) / e2e_s,
            "end_to_end": "CNF list -> packed clauses -> job list -> upload -> launch -> verdicts on the host "
                          "(median of 5)",
            "end_to_end_streamed": stream,
            "host_s": host,
            "branch_nodes": nodes_total,
            "leaves": int(js["leaves"].sum()),
            **extra,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC.get(("search", width, len(cnfs)), (None, None))[0],
            "traffic_source": MEASURED_TRAFFIC.get(("search", width, len(cnfs)), (None, None))[1],
            "kernel": "k_search",
            "bytes_per_step": alg_bytes,
            "bytes_per_launch_avg": alg_bytes / max(1, solver.rounds),
            "launch_avg_us": launch_avg_us,
            "binding_limit": "latency: the longest job's chain of dependent branch nodes (each a few "
                             "propagation rounds of ballots and scalar unit steps), not HBM",
            "critical_path": {"busiest_job_nodes": int(js["nodes"][busiest]),
                              "busiest_job_conflicts": int(js["conflicts"][busiest]),
                              "mean_job_nodes": float(js["nodes"].mean()),
                              "us_per_node_on_busiest": kernel_ms * 1000.0 / max(1, int(js["nodes"][busiest]
                                                                                    + js["conflicts"][busiest]))},
        },
        "cpu_baseline": cpu,
        "parity": parity,
        "configs": nested,
    }
    print(json.dumps(out), flush=True)
    bad = [parity[k] for k in ("oracle_search_mismatches", "reference_path_mismatches", "dpll_mismatches",
                               "end_to_end_mismatches", "streamed_mismatches", "undecided") if parity.get(k)]
    nested_bad = any(v.get("status") == "mismatch" for v in nested.values())
    return 1 if bad or nested_bad else 0


def bench_batched_host(args) -> int:
    """The earlier batched pipeline: host cube planning to width W, then one persistent
    contraction launch over device-resident jobs (kernel time only in ``value``)."""
    import torch

    from tensorsat_amd.solver import BatchSolver
    from tensorsat_amd.workloads import batched_3sat

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    cnfs = batched_3sat(args.instances)
    log(f"generated {len(cnfs)} instances in {time.perf_counter() - t0:.1f}s")
    tune = None
    if args.width == "auto":
        from tensorsat_amd.solver import autotune_width

        t0 = time.perf_counter()
        width, table = autotune_width(cnfs, first_pass=args.first_pass, mode=args.mode, threads=args.threads)
        tune = {"chosen": width, "seconds": time.perf_counter() - t0, "sample": 512,
                "candidates": {str(k): v for k, v in table.items()}}
        log(f"autotuned cube width W={width} in {tune['seconds']:.1f}s: " +
            ", ".join(f"W={k}: plan {v['plan_s']:.3f}s + gpu {v['gpu_ms']:.3f}ms" for k, v in table.items()))
    else:
        width = int(args.width)
    solver = BatchSolver(width=width, first_pass=args.first_pass, mode=args.mode, threads=args.threads,
                         num_workgroups=args.workgroups or None,
                         lds={"auto": None, "lds": True, "global": False}[args.kernel],
                         round_bounds=[int(x) for x in args.rounds.split(",")] + [1 << 30] if args.rounds else None)
    t0 = time.perf_counter()
    solver.prepare(cnfs)
    prep_s = time.perf_counter() - t0
    st = solver.stats
    log(f"plan {st.plan_s:.2f}s: pass1 {st.pass1_jobs} jobs, pass2 {st.pass2_jobs} jobs for {st.pass2_instances} "
        f"instances; {st.host_unsat} decided UNSAT on the host")
    for _ in range(args.warmup):
        solver.run()
    torch.cuda.synchronize()
    # Eager launches of the same steps from the host, for reference.
    t_start = time.perf_counter()
    for k in range(args.steps):
        solver.launch()
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t_start) * 1000.0 / args.steps
    # The timed steps: one HIP graph holding all K steps (each the reset fill and the
    # persistent launch of the whole batch), replayed once; a graph of one step is also
    # timed (its launch costs more host time than the step takes on the device).
    solver.capture(steps=1)
    for _ in range(args.warmup):
        solver.replay()
    torch.cuda.synchronize()
    t_g = time.perf_counter()
    for k in range(args.steps):
        solver.replay()
    torch.cuda.synchronize()
    graph1_ms = (time.perf_counter() - t_g) * 1000.0 / args.steps
    solver.capture(steps=args.steps)
    solver.replay()  # warm
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    solver.replay()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    # Device time per step: events around each step's kernels (a loop of its own, so the
    # events do not stretch the timed steps).
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for k in range(args.steps):
        solver.launch(events=ev[k])
    torch.cuda.synchronize()
    kernel_ms = [a.elapsed_time(b) for a, b in ev]
    verdicts = solver.verdicts()
    ms_per_step = wall * 1000.0 / args.steps
    avg_kernel_ms = float(np.mean(kernel_ms))
    value = len(cnfs) / (ms_per_step / 1000.0)

    bytes_per_launch = 0.0
    jobs_executed = 0
    for prog, mask in zip(solver.progs, solver.executed_jobs()):
        bytes_per_launch += float(job_bytes(prog)[mask].sum())
        jobs_executed += int(mask.sum())
    achieved = bytes_per_launch / (avg_kernel_ms / 1000.0) / 1e9

    cpu = None
    parity = {"checked": 0, "mismatches": 0}
    if not args.no_cpu:
        ref, dt = cpu_baseline(solver, len(cnfs), args.mode, args.cpu_seconds)
        n_done = len(ref)
        mism = sum(int(verdicts[i] != v) for i, v in ref.items())
        parity = {"checked": n_done, "mismatches": mism}
        cpu = {
            "value": n_done / dt,
            "unit": "instances/s",
            "cores": 1,
            "kind": "port",
            "sample": f"first {n_done} of the {len(cnfs)} instances, same cubes and trees, numpy step loop "
                      f"(oracle/executor.py) with early exit, {dt:.1f}s, single thread, {platform.processor() or 'host CPU'}",
        }
        log(f"cpu baseline {cpu['value']:.1f} inst/s on {n_done} instances; parity mismatches {mism}")
        from oracle import dpll

        dp = np.array([dpll.solve(c.num_vars, c.clauses) for c in cnfs], dtype=np.uint8)
        parity["dpll_checked"] = len(cnfs)
        parity["dpll_mismatches"] = int((dp != verdicts).sum())
        log(f"independent DPLL cross-check: {parity['dpll_mismatches']} mismatches over {len(cnfs)}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": "batched random 3-SAT decisions: 4096 instances, n~U{20..80}, ratio 4.26/3.5 "
                        "(default_rng(10000+i)), conditioned slicing to cube width W, device-resident jobs",
            "instances": len(cnfs),
            "cube_width": width,
            "autotune": tune,
            "host_decided_unsat": st.host_unsat,
            "first_pass_cubes": args.first_pass,
            "jobs": st.pass1_jobs + st.pass2_jobs,
            "jobs_executed": jobs_executed,
            "pass2_instances": st.pass2_instances,
            "mode": args.mode,
            "workgroups": solver.execs[0].num_workgroups,
            "kernel": batch_kernel_name(solver.execs[0]),
            "rounds_per_pass": [len(ex.rounds) for ex in solver.execs],
            "plan_s": st.plan_s,
            "end_to_end_s": st.plan_s + ms_per_step / 1000.0,
            "sat": int(verdicts.sum()),
            "kernel_ms": avg_kernel_ms,
            "launch": "one hipGraph replay holding all timed steps (each: reset fill + one persistent launch)",
            "eager_ms_per_step": eager_ms,
            "graph_one_step_ms": graph1_ms,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC.get(("batched", width, len(cnfs), args.first_pass), (None, None))[0],
            "traffic_source": MEASURED_TRAFFIC.get(("batched", width, len(cnfs), args.first_pass), (None, None))[1],
            "kernel": batch_kernel_name(solver.execs[0]),
            "bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(out), flush=True)
    return 0 if parity["mismatches"] == 0 and parity.get("dpll_mismatches", 0) == 0 else 1


def _step_bytes(steps, elem: int = 1) -> float:
    """Algorithmic bytes of a step list: 2^rank(lhs) + 2^rank(rhs) + 2^rank(out) entries."""
    norhs = (steps["flags"] & 2) != 0
    b = np.ldexp(1.0, steps["lhs_rank"].astype(np.int64)) + np.ldexp(1.0, steps["out_rank"].astype(np.int64))
    b += np.where(norhs, 0.0, np.ldexp(1.0, steps["rhs_rank"].astype(np.int64)))
    return float(b.sum()) * elem


def run_single(args, seeds, log) -> dict:
    """
    One large random 3-SAT instance on the whole GPU (n=120, m=510, seeds 0..): conditioned
    slicing into cubes of width up to 26 (tensors up to 2^26 bytes), cubes in depth-first
    order in gangs of 8 advanced together (``ts_program_run_gang``: the runs of
    small steps of a gang's cubes in one launch, one workgroup per cube, their large steps
    in one grid-wide launch; ``single.decide``), until the first satisfiable cube.  Reported: plan and GPU seconds per instance, and the step
    kernels' algorithmic bytes over the GPU time (HBM bound).  Verdicts are checked against
    the independent DPLL.  CPU baseline: the oracle's numpy step loop on the same cubes of
    the first seed, one cube per process on all host cores, for a bounded wall time,
    scaled to the cubes the GPU ran.
    """
    import torch

    from tensorsat_amd import program as PG
    from tensorsat_amd.executor import DeviceProgram, run_job
    from tensorsat_amd.workloads import single_3sat

    from tensorsat_amd.single import decide

    gang = args.gang  # None: single.decide's choice (8 for cubes of different shapes)
    rows = []
    tot_bytes = tot_gpu = tot_plan = 0.0
    mism = 0
    progs = []
    for seed in seeds:
        cnf = single_3sat(seed)
        t0 = time.perf_counter()
        prog = PG.sat_batch([cnf], width=26, plan_trials=2)
        plan_s = time.perf_counter() - t0
        progs.append(prog)
        kw = dict(program=prog, mode=args.mode, streams=args.streams, gang=gang)
        decide(cnf, **kw)  # warm-up: code objects, arenas
        r = decide(cnf, **kw)
        gpu_s = r.gpu_s
        launched = r.stats.get("cubes_launched", r.cubes_run)
        nbytes = float(sum(_step_bytes(prog.job_steps(j)) for j in range(launched)))
        want = None
        if not args.no_cpu:  # CPU checking leg: the independent DPLL
            from oracle import dpll

            want = int(dpll.solve(cnf.num_vars, cnf.clauses))
            mism += int(r.value != want)
        rows.append({"seed": seed, "verdict": r.value, "dpll": want, "plan_s": plan_s, "gpu_s": gpu_s,
                     "cubes": len(prog.jobs), "cubes_run": r.cubes_run, "cubes_launched": launched,
                     "max_rank": int(prog.stats[0, 4]), "bytes": nbytes})
        tot_bytes += nbytes
        tot_gpu += gpu_s
        tot_plan += plan_s
        log(f"n=120 seed {seed}: verdict {r.value} (dpll {want}), plan {plan_s:.2f}s, gpu {gpu_s:.4f}s, "
            f"{r.cubes_run}/{len(prog.jobs)} cubes ({launched} launched), {nbytes / gpu_s / 1e9:.0f} GB/s")
    gbs = tot_bytes / tot_gpu / 1e9
    cpu = None
    if not args.no_cpu:
        cores = host_cores()
        prog, ran = progs[0], rows[0]["cubes_run"]
        k = min(ran, cores)
        t0 = time.perf_counter()
        with _pool(prog, cores) as ex:
            list(ex.map(_pool_job, range(k), [args.mode] * k))
        dt = time.perf_counter() - t0
        per_inst = dt * ran / k
        cpu = {"value": per_inst, "unit": "s per instance", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
               "sample": f"seed {rows[0]['seed']}: the first {k} of the {ran} cubes the GPU ran, through the "
                         f"oracle's numpy step loop in {cores} processes, {dt:.1f}s wall, scaled by {ran}/{k}"}
        log(f"n=120 cpu: {k} cubes in {dt:.1f}s -> {per_inst:.1f}s per instance")
    return {
        "metric": "plan+contract s per instance, single n=120 3-SAT decision (bit-exact)",
        "value": (tot_plan + tot_gpu) / len(rows), "unit": "s", "steps": len(rows),
        "ms_per_step": tot_gpu * 1000.0 / len(rows), "higher_is_better": False, "dtype": "u8",
        "config": {"workload": "single random 3-SAT n=120 m=510 (seeds 0..), cubes of width <= 26 in gangs of up to "
                               f"{gang or 8} advanced together (0: one at a time on {args.streams} streams)",
                   "plan_s_mean": tot_plan / len(rows), "gpu_s_mean": tot_gpu / len(rows), "rows": rows},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "k_gang_step / k_gang_seq: algorithmic bytes of every step of the "
                                                "launched cubes over the decision's wall time"},
        "cpu_baseline": cpu,
        "parity": {"checked": len(rows), "dpll_mismatches": mism},
        "status": "mismatch" if mism else "ok",
    }


def run_plain(args, seeds, log) -> dict:
    """
    Config 1, the reference's own call shape on one GPU: a random 3-SAT instance with n=40,
    m=170 (``workloads.plain_cpu_3sat``), its flattened bintree network
    (``(cnf.inputs(None) >> cnf.diagram()).flatten()``) contracted along a greedy tree
    (``PlannedContraction(..., optimize="greedy")``, ``examples/test_sat_solver.py:5-9``)
    sliced to intermediates of at most 2^26 entries, every slice compiled into one job of
    step records (contracted wires innermost) and the slices decided in gangs advanced
    together with early exit (``single.decide``).  Verdicts against the independent DPLL.  CPU: the
    oracle's numpy step loop on the same jobs on all host cores, sampled and scaled.
    """
    from tensorsat_amd import program as PG
    from tensorsat_amd.contractions import PlannedContraction
    from tensorsat_amd.rel import DeviceRel
    from tensorsat_amd.single import decide
    from tensorsat_amd.workloads import plain_cpu_3sat

    gang = args.gang  # None: single.decide's choice (32 for slices of one structure)
    rows, progs = [], []
    mism = 0
    for seed in seeds:
        cnf = plain_cpu_3sat(seed)
        d = (cnf.inputs(None) >> cnf.diagram()).flatten()
        t0 = time.perf_counter()
        c = PlannedContraction(DeviceRel, d.wiring, optimize="greedy", slicing_target=26)
        plan_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        prog = c.contract_program(d)
        compile_s = time.perf_counter() - t0
        progs.append(prog)
        kw = dict(program=prog, mode=args.mode, streams=args.streams, gang=gang)
        decide(cnf, **kw)  # warm-up
        r = decide(cnf, **kw)
        launched = r.stats.get("cubes_launched", r.cubes_run)
        st = prog.steps[: int(prog.jobs["step_begin"][launched - 1] + prog.jobs["num_steps"][launched - 1])]
        dense, ops, nbytes = PG.step_work(st, args.min_sum_bits)
        want = None
        if not args.no_cpu:
            from oracle import dpll

            want = int(dpll.solve(cnf.num_vars, cnf.clauses))
            mism += int(r.value != want)
        rows.append({"seed": seed, "verdict": r.value, "dpll": want, "plan_s": plan_s, "compile_s": compile_s,
                     "gpu_s": r.gpu_s, "slices": len(prog.jobs), "slices_run": r.cubes_run,
                     "slices_launched": launched, "sliced_wires": len(c.tree.sliced),
                     "log2_max_size": float(c.tree.cost.log2_max_size), "ops": float(ops.sum()),
                     "bytes": float(nbytes.sum()), "mfma_op_share": float(ops[dense].sum() / ops.sum())})
        log(f"n=40 seed {seed}: verdict {r.value} (dpll {want}), plan {plan_s:.2f}s + compile {compile_s:.2f}s, "
            f"gpu {r.gpu_s:.4f}s, {r.cubes_run}/{len(prog.jobs)} slices ({launched} launched), "
            f"{nbytes.sum() / r.gpu_s / 1e9:.0f} GB/s, MFMA-routed ops {ops[dense].sum() / ops.sum():.1%}")
    gpu = float(np.mean([r["gpu_s"] for r in rows]))
    host = float(np.mean([r["plan_s"] + r["compile_s"] for r in rows]))
    gbs = sum(r["bytes"] for r in rows) / sum(r["gpu_s"] for r in rows) / 1e9
    cpu = None
    if not args.no_cpu:
        cores = host_cores()
        prog, ran = progs[0], rows[0]["slices_run"]
        k = min(ran, cores)
        t0 = time.perf_counter()
        with _pool(prog, cores) as ex:
            list(ex.map(_pool_job, range(k), [args.mode] * k))
        dt = time.perf_counter() - t0
        per_inst = dt * ran / k
        cpu = {"value": per_inst + rows[0]["plan_s"] + rows[0]["compile_s"], "unit": "s per instance",
               "cores": cores, "kind": "port", "cpu_model": cpu_model(),
               "sample": f"seed {rows[0]['seed']}: {k} of the {ran} slices the GPU ran, through the oracle's numpy "
                         f"step loop in {cores} processes, {dt:.1f}s wall, scaled by {ran}/{k}, plus the same "
                         f"plan and compile time"}
        log(f"n=40 cpu: {k} slices in {dt:.1f}s -> {per_inst:.1f}s per instance")
    return {
        "metric": "plan+contract s per instance, plain n=40 3-SAT decision through the reference's call shape "
                  "(greedy tree, bit-exact)",
        "value": host + gpu, "unit": "s", "steps": len(rows), "ms_per_step": gpu * 1000.0,
        "higher_is_better": False, "dtype": "u8",
        "config": {"workload": "plain random 3-SAT n=40 m=170 (seeds 0..), bintree network, greedy tree sliced to "
                               f"2^26 entries, one compiled job per slice, gangs of up to {gang or 32} slices, "
                               "early exit",
                   "host_s_mean": host, "gpu_s_mean": gpu, "rows": rows},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "k_gang_step / k_gang_seq: algorithmic bytes of every step of the "
                                                "launched slices over the decision's wall time"},
        "cpu_baseline": cpu,
        "parity": {"checked": len(rows), "dpll_mismatches": mism},
        "status": "mismatch" if mism else "ok",
    }


def bench_plain(args) -> int:
    out = run_plain(args, range(args.seeds), lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
    out.update({"n_gpus": 1, "warmup": 1, "scaling": "weak", "vs_baseline": None, "data": "synthetic"})
    print(json.dumps(out), flush=True)
    return 1 if out["status"] == "mismatch" else 0


def run_gnn(args, log, batch: int = 32, steps: int = 10) -> dict:
    """
    The learned planner network's training step (6 message-passing layers of width 128,
    bf16 operands, fp32 accumulation; ``learn.model.GNN``) on one batch of 32 training
    instances of the config-5 distribution (``learn.data.ksat``, seeds 10^6..): forward,
    backward and Adam, all ``ts_gnn_*`` kernels.  Roofline of the GEMM kernel (``k_gemm``,
    v_mfma_f32_16x16x32_bf16) over the step's GEMMs, each timed with HIP events on the launch
    stream: their products are skinny (N = 128, K = 384 or the row count), 97 flops per byte
    at most against the 312 of the bf16 ridge (2.5 PFLOP/s over 8 TB/s), so the bound is
    HBM: achieved = algorithmic bytes (operands once, output once, twice when accumulated)
    over the summed durations; the flop rate is reported beside it.
    """
    import torch

    from tensorsat_amd.learn import data as D
    from tensorsat_amd.learn import graph as G
    from tensorsat_amd.learn.model import GNN

    g = G.build([D.ksat(10**6 + i) for i in range(batch)])
    dg = G.DeviceGraph(g)
    target = torch.rand(dg.V, device="cuda")
    model = GNN(seed=0)
    for _ in range(2):  # warm-up
        model.forward(dg, target=target, train=True)
        model.adam_step(lr=1e-4)
    torch.cuda.synchronize()
    model.gemm_log = []
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        model.forward(dg, target=target, train=True)
        model.adam_step(lr=1e-4)
    e1.record()
    torch.cuda.synchronize()
    step_ms = e0.elapsed_time(e1) / steps
    flops = sum(f for f, _, _, _ in model.gemm_log)
    nbytes = sum(b for _, b, _, _ in model.gemm_log)
    gemm_s = sum(a.elapsed_time(b) for _, _, a, b in model.gemm_log) / 1000.0
    calls = len(model.gemm_log) // steps
    model.gemm_log = None
    tflops = flops / gemm_s / 1e12
    gbs = nbytes / gemm_s / 1e9
    log(f"gnn: {dg.V} variables, {dg.C} clauses, step {step_ms:.2f} ms, {calls} GEMMs per step, "
        f"GEMM {gemm_s / steps * 1e3:.2f} ms per step at {gbs:.0f} GB/s, {tflops:.1f} TFLOP/s")
    return {
        "metric": "ms per training step, learned planner network (batch of 32 instances)",
        "value": step_ms, "unit": "ms", "steps": steps, "ms_per_step": step_ms, "higher_is_better": False,
        "dtype": "bf16",
        "config": {"workload": f"message-passing network, 6 layers x 128, batch of {batch} config-5 training "
                               f"instances ({dg.V} variables, {dg.C} clauses), forward + backward + Adam",
                   "gemm_calls_per_step": calls, "gemm_ms_per_step": gemm_s / steps * 1e3,
                   "gemm_tflops": tflops, "gemm_tflops_frac_of_bf16_peak": tflops / 2500.0,
                   "gemm_flops_per_byte": flops / nbytes},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "k_gemm: algorithmic bytes of the step's GEMMs over their summed "
                                                "durations"},
        "cpu_baseline": None,
        "status": "ok",
    }


def bench_single(args) -> int:
    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    line = run_single(args, range(args.seeds), log)
    line.update({"n_gpus": 1, "warmup": 0, "scaling": "weak", "vs_baseline": None, "data": "synthetic"})
    print(json.dumps(line), flush=True)
    return 0 if line["status"] == "ok" else 1


def run_count(args, seeds, log) -> dict:
    """
    Exact model counts (#SAT) of random 3-SAT n=100, m=300 (seeds 0..): cubes of width 18
    counted in uint64 by the persistent counting kernel, each cube's result times
    2^free added per instance with overflow detection.  Parity: the independent model
    counter's count (``oracle/sharp.c``, committed in tests/golden for seeds 0..9) and a
    second decomposition (width 16).  CPU baseline: the oracle's counting step loop on a
    bounded sample of the same cubes, one cube per process on all host cores, scaled to
    all cubes.
    """
    import torch

    from tensorsat_amd import program as PG
    from tensorsat_amd.executor import BatchCounter
    from tensorsat_amd.workloads import count_3sat

    golden = {}
    gpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "sharp_n100_counts.json")
    if os.path.exists(gpath):
        golden = {int(k): int(v) for k, v in json.load(open(gpath))["counts"].items()}
    rows = []
    mism = 0
    tot_gpu = tot_plan = tot_bytes = 0.0
    progs = []
    for seed in seeds:
        cnf = count_3sat(seed)
        t0 = time.perf_counter()
        prog = PG.sat_batch([cnf], width=18, plan_trials=1)
        plan_s = time.perf_counter() - t0
        progs.append(prog)
        bc = BatchCounter(prog)
        bc.run()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        bc.launch()
        b.record()
        torch.cuda.synchronize()
        gpu_s = a.elapsed_time(b) / 1000.0
        (c18,) = bc.run()
        (c16,) = BatchCounter(PG.sat_batch([cnf], width=16, plan_trials=1)).run()
        want = golden.get(seed)
        mism += int(c18 != c16) + int(want is not None and c18 != want)
        nbytes = _step_bytes(prog.steps, 8)
        rows.append({"seed": seed, "count": str(c18), "count_w16": str(c16),
                     "independent_count": None if want is None else str(want), "plan_s": plan_s, "gpu_s": gpu_s,
                     "cubes": len(prog.jobs), "bytes": nbytes})
        tot_gpu += gpu_s
        tot_plan += plan_s
        tot_bytes += nbytes
        log(f"n=100 seed {seed}: count {c18} (width 16: {c16}, independent: {want}), plan {plan_s:.2f}s, "
            f"gpu {gpu_s:.3f}s, {len(prog.jobs)} cubes")
    gbs = tot_bytes / tot_gpu / 1e9
    cpu = None
    if not args.no_cpu:
        from concurrent.futures import wait

        cores = host_cores()
        prog = progs[0]
        nj = len(prog.jobs)
        sample = np.random.default_rng(0).permutation(nj)[: min(nj, 64 * cores)]
        t0 = time.perf_counter()
        with _pool(prog, cores) as ex:
            wait([ex.submit(_pool_job, int(j), "COUNT") for j in sample])
        dt = time.perf_counter() - t0
        per_inst = dt * nj / len(sample)
        cpu = {"value": per_inst, "unit": "s per instance", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
               "sample": f"seed {rows[0]['seed']}: {len(sample)} random cubes of {nj} through the oracle's "
                         f"counting step loop in {cores} processes, {dt:.1f}s wall, scaled by {nj}/{len(sample)}"}
        log(f"n=100 cpu: {len(sample)} cubes in {dt:.1f}s -> {per_inst:.1f}s per instance")
    return {
        "metric": "plan+contract s per instance, exact #SAT n=100 m=300 (uint64, checked)",
        "value": (tot_plan + tot_gpu) / len(rows), "unit": "s", "steps": len(rows),
        "ms_per_step": tot_gpu * 1000.0 / len(rows), "higher_is_better": False, "dtype": "u64",
        "config": {"workload": "#SAT of random 3-SAT n=100 m=300 (seeds 0..), cubes of width <= 18, "
                               "persistent uint64 counting kernel",
                   "plan_s_mean": tot_plan / len(rows), "gpu_s_mean": tot_gpu / len(rows), "rows": rows},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "k_batch_count"},
        "cpu_baseline": cpu,
        "parity": {"checked": len(rows), "mismatches": mism},
        "status": "mismatch" if mism else "ok",
    }


def bench_count(args) -> int:
    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    line = run_count(args, range(args.seeds), log)
    line.update({"n_gpus": 1, "warmup": 1, "scaling": "weak", "vs_baseline": None, "data": "synthetic"})
    print(json.dumps(line), flush=True)
    return 0 if line["status"] == "ok" else 1


def bench_rewire(args) -> int:
    """
    Kernels (a) and (b) of ts_rewire on a 2^27-byte relation (the largest device rank): a
    random permutation of its 27 bit axes (LDS-staged tiles, 2 x 2^27 bytes moved) and an
    existential reduction over 11 of them (wave per out entry, ballot OR; 2^27 + 2^16
    bytes).  GB/s of algorithmic bytes
    against the 8 TB/s HBM peak; a plain device copy of the same bytes is timed beside the
    permutation.  Parity: sampled entries against numpy on the host.
    """
    import torch

    from tensorsat_amd import _native as N

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    r = 27
    g = torch.Generator(device="cuda").manual_seed(3)
    x = (torch.rand(1 << r, device="cuda", generator=g) < 0.5).to(torch.uint8)
    rng = np.random.default_rng(5)
    perm = [int(p) for p in rng.permutation(r)]
    keep = sorted(int(p) for p in rng.choice(r, size=16, replace=False))
    out_p = torch.empty(1 << r, dtype=torch.uint8, device="cuda")
    out_r = torch.empty(1 << 16, dtype=torch.uint8, device="cuda")
    L = N.lib()

    def rw(out, ports):
        N.check(L.ts_rewire(N.dptr(x), r, N.dptr(out), len(ports), N.i32arr(ports), N.MODES[args.mode],
                            N.stream_ptr()), "ts_rewire")

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.steps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / args.steps

    ms_p = timed(lambda: rw(out_p, perm))
    ms_r = timed(lambda: rw(out_r, keep))
    scratch = torch.empty_like(x)
    ms_t = timed(lambda: scratch.copy_(x))  # a plain copy of the same bytes: the ideal
    gbs_p = 2.0 * (1 << r) / (ms_p / 1000.0) / 1e9
    gbs_r = ((1 << r) + (1 << 16)) / (ms_r / 1000.0) / 1e9
    gbs_t = 2.0 * (1 << r) / (ms_t / 1000.0) / 1e9
    # parity on sampled entries
    xh = x.cpu().numpy().reshape((2,) * r)
    ph = out_p.cpu().numpy()
    idx = rng.integers(0, 1 << r, size=4096)
    bits = ((idx[:, None] >> np.arange(r - 1, -1, -1)) & 1)  # out axis j = bit r-1-j
    src = np.zeros((len(idx), r), np.int64)
    for j, p in enumerate(perm):
        src[:, p] = bits[:, j]
    mism = int((ph[idx] != xh[tuple(src.T)]).sum())
    red = np.asarray(xh).reshape(-1)
    rh = out_r.cpu().numpy()
    ridx = rng.integers(0, 1 << 16, size=64)
    drop = [p for p in range(r) if p not in keep]
    for o in ridx:
        sl = [slice(None)] * r
        for j, p in enumerate(keep):
            sl[p] = (int(o) >> (15 - j)) & 1
        cnt = int(xh[tuple(sl)].sum())
        want = (cnt > 0) if args.mode == "GT0" else (cnt % 256 != 0)
        mism += int(rh[o] != want)
    log(f"permute {ms_p:.3f} ms ({gbs_p:.0f} GB/s; plain copy {ms_t:.3f} ms, {gbs_t:.0f} GB/s), "
        f"reduce {ms_r:.3f} ms ({gbs_r:.0f} GB/s); sampled mismatches {mism}")
    print(json.dumps({
        "metric": "GB/s rewire kernels on a 2^27-byte relation (permutation; reduction over 11 axes)",
        "value": gbs_p, "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_p, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "ts_rewire on a rank-27 0/1 tensor: random axis permutation, and keep 16 of 27 axes",
                   "permute_ms": ms_p, "permute_gbs": gbs_p, "copy_ms": ms_t, "copy_gbs": gbs_t,
                   "reduce_ms": ms_r, "reduce_gbs": gbs_r, "mode": args.mode},
        "roofline": {"bound": "hbm", "achieved": gbs_p, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs_p / HBM_PEAK_GBS, "traffic": None, "kernel": "k_permute_lds",
                     "reduce": {"achieved": gbs_r, "frac": gbs_r / HBM_PEAK_GBS, "kernel": "k_reduce"}},
        "cpu_baseline": None,
        "parity": {"sampled": 4096 + 64, "mismatches": mism},
    }), flush=True)
    return 0 if mism == 0 else 1


def bench_copy(args) -> int:
    """
    Calibration for hardware traffic counters: a plain device-to-device copy of 1 GiB
    (torch ``copy_``), --steps times.  Under ``rocprofv3 --pmc FETCH_SIZE`` (then
    ``WRITE_SIZE``) each dispatch should read (write) 2^20 KiB; the measured ratio scales
    the counters of the other kernels.
    """
    import torch

    n = 1 << 30
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    gbs = 2.0 * n / (ms / 1000.0) / 1e9
    print(json.dumps({"metric": "GB/s device copy (counter calibration)", "value": gbs, "unit": "GB/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
                      "config": {"workload": "copy of 1 GiB", "bytes_read": n, "bytes_written": n},
                      "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": gbs / HBM_PEAK_GBS, "traffic": 2.0 * n, "kernel": "copy"},
                      "cpu_baseline": None}), flush=True)
    return 0


def bench_learned(args) -> int:
    """
    Learned planner vs plain planner, end to end (host planning + GPU contraction), on the
    learned-planner workload: k ~ U{3,4,5}, n ~ U[150, 300], m = 0.9 x threshold x n
    (``workloads.learned_ksat`` with seeds 3e6 + i, disjoint from the training seeds).
    Each instance is decided twice with the same budget: plain (``single.decide_budget``:
    multi-trial planner trees, dynamic branching) and learned (``single.decide_learned``:
    network scores from one batched GPU inference steer the trees and the branching order,
    with the plain search as fallback in alternating time slices).  Reported: time to
    solution of each, solved counts, and verdict identity where both decide.
    """
    import torch

    from tensorsat_amd.learn.propose import LearnedPlanner
    from tensorsat_amd.single import decide_budget, decide_learned
    from tensorsat_amd.workloads import learned_ksat

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    cnfs, seeds = [], []
    i = 0
    while len(cnfs) < args.learned_instances:
        c = learned_ksat(i, base=3_000_000)
        if args.learned_k == 0 or max(len(cl) for cl in c.clauses) == args.learned_k:
            cnfs.append(c)
            seeds.append(3_000_000 + i)
        i += 1
    lp = LearnedPlanner()
    lp.scores(cnfs[:1])  # warm-up (kernel load)
    scores = lp.scores(cnfs)
    infer_s = lp.last_infer_s
    log(f"network inference for {len(cnfs)} instances: {infer_s * 1000:.1f} ms")
    rows = []
    for i, cnf in enumerate(cnfs):
        k = max(len(c) for c in cnf.clauses)
        r = {"seed": seeds[i], "k": k, "n": cnf.num_vars, "m": len(cnf.clauses)}
        torch.cuda.synchronize()
        for name in ("plain", "learned"):
            t0 = time.perf_counter()
            if name == "plain":
                res = decide_budget(cnf, width=args.learned_width, budget_s=args.budget)
            else:
                res = decide_learned(cnf, scores[i], width=args.learned_width, budget_s=args.budget)
            dt = time.perf_counter() - t0 + (infer_s / len(cnfs) if name == "learned" else 0.0)
            r[name] = {"verdict": res.value, "s": dt, "plan_s": res.plan_s, "gpu_s": res.gpu_s, "cubes": res.cubes,
                       "decided": res.stats.get("decided", "budget"),
                       "learned_trees": res.stats["learned_trees"], "fallback_trees": res.stats["fallback_trees"]}
        log(f"seed={seeds[i]} k={k} n={cnf.num_vars}: plain {r['plain']['verdict']} in {r['plain']['s']:.2f}s "
            f"({r['plain']['cubes']} cubes), learned {r['learned']['verdict']} in {r['learned']['s']:.2f}s "
            f"({r['learned']['cubes']} cubes, trees {r['learned']['learned_trees']} learned / "
            f"{r['learned']['fallback_trees']} fallback)")
        rows.append(r)
    both = [r for r in rows if r["plain"]["verdict"] >= 0 and r["learned"]["verdict"] >= 0]
    mism = sum(int(r["plain"]["verdict"] != r["learned"]["verdict"]) for r in both)
    tp = sum(r["plain"]["s"] if r["plain"]["verdict"] >= 0 else args.budget for r in rows)
    tl = sum(r["learned"]["s"] if r["learned"]["verdict"] >= 0 else args.budget for r in rows)
    solved_p = sum(r["plain"]["verdict"] >= 0 for r in rows)
    solved_l = sum(r["learned"]["verdict"] >= 0 for r in rows)
    speed = [r["plain"]["s"] / r["learned"]["s"] for r in both]
    line = {
        "metric": METRIC, "value": solved_l, "unit": "instances solved within budget (learned planner)",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": tl * 1000.0 / max(len(rows), 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "learned planner vs plain planner: random k-SAT, k~U{3,4,5}, n~U[150,300], "
                               "0.9 x threshold, seeds 3e6+i, budgeted cube search + GPU contraction",
                   "instances": len(rows), "k_filter": args.learned_k or None, "budget_s": args.budget, "cube_width": args.learned_width,
                   "solved_plain": solved_p, "solved_learned": solved_l, "both_solved": len(both),
                   "verdict_mismatches": mism, "total_s_plain": tp, "total_s_learned": tl,
                   "geomean_speedup_both_solved": float(np.exp(np.mean(np.log(speed)))) if speed else None,
                   "inference_ms": infer_s * 1000.0, "rows": rows},
        "roofline": None, "cpu_baseline": None,
    }
    torch.cuda.synchronize()
    print(json.dumps(line), flush=True)
    return 0 if mism == 0 else 1


def bench_dense(args) -> int:
    """
    One dense Boolean step: the relational composition out[m, n] = OR_k L[m, k] & R[n, k]
    with 2^M rows, 2^N columns and 2^K contracted entries, in the compiled layout (summed
    wires innermost in both operands), routed to the int8 MFMA kernel.  Operand density
    2^(-K/2) keeps both output values frequent.  Algorithmic work: 2 * 2^(M+N+K) int8
    ops; bytes 2^(M+K) + 2^(N+K) + 2^(M+N).  The VALU route is timed on the same step, and
    rows of the result are checked against exact CPU counts.
    """
    import torch

    from tensorsat_amd import _native as N

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    mb, nb, kb = args.dense_bits
    K = list(range(kb))[::-1]
    lw = [100 + i for i in range(mb)] + K
    rw = [200 + i for i in range(nb)] + K
    ow = [100 + i for i in range(mb)] + [200 + i for i in range(nb)]
    g = torch.Generator(device="cuda").manual_seed(1)
    dens = 2.0 ** (-kb / 2)
    lhs = (torch.rand(1 << (mb + kb), device="cuda", generator=g) < dens).to(torch.uint8)
    rhs = (torch.rand(1 << (nb + kb), device="cuda", generator=g) < dens).to(torch.uint8)
    out = torch.empty(1 << (mb + nb), dtype=torch.uint8, device="cuda")
    mode = N.MODES[args.mode]
    L = N.lib()

    def step():
        N.check(L.ts_contract2(N.dptr(lhs), len(lw), N.i32arr(lw), N.dptr(rhs), len(rw), N.i32arr(rw), N.dptr(out),
                               len(ow), N.i32arr(ow), mode, N.stream_ptr()), "ts_contract2")

    def timed(steps):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        a.record()
        for _ in range(steps):
            step()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps, (time.perf_counter() - t0) * 1000.0 / steps

    # Tile autotuner: every MFMA workgroup tile timed on this step shape, the fastest
    # registered for it (counts are exact with any tile).
    from tensorsat_amd import dense_tune as DT

    tune = DT.autotune_tiles([(mb, nb, 0, kb)], reps=3).get((mb, nb, 0, kb))
    log(f"tile autotune: {tune}")
    tile = DT.tile_of(mb, nb, 0, kb)
    kernel_ms, ms_per_step = timed(args.steps)
    res = out.cpu().numpy().reshape(1 << mb, 1 << nb)
    N.check(L.ts_dense_config(0, 6), "ts_dense_config")
    try:
        valu_ms, _ = timed(1)
        valu = out.cpu().numpy().reshape(1 << mb, 1 << nb)
    finally:
        N.check(L.ts_dense_config(1, 6), "ts_dense_config")
    # Exact counts for a sample of rows (float32 products of 0/1 are exact below 2^24).
    rows = np.random.default_rng(0).choice(1 << mb, size=min(64, 1 << mb), replace=False)
    Lh = lhs.cpu().numpy().reshape(1 << mb, 1 << kb)
    Rh = rhs.cpu().numpy().reshape(1 << nb, 1 << kb)
    cnt = Lh[rows].astype(np.float32) @ Rh.astype(np.float32).T
    want = (cnt > 0) if args.mode == "GT0" else ((cnt.astype(np.int64) % 256) != 0)
    mism = int((res[rows] != want).sum())
    valu_mism = int((valu != res).sum())
    ops = 2.0 * 2.0 ** (mb + nb + kb)
    nbytes = 2.0 ** (mb + kb) + 2.0 ** (nb + kb) + 2.0 ** (mb + nb)
    tops = ops / (kernel_ms / 1000.0) / 1e12
    log(f"MFMA {kernel_ms:.3f} ms ({tops:.1f} TOP/s), VALU {valu_ms:.3f} ms; sampled-row mismatches {mism}, "
        f"MFMA vs VALU mismatches {valu_mism}; ones {float(res.mean()):.3f}")
    cpu = None
    if not args.no_cpu:
        from oracle import finrel as F

        # The reference-equivalent step (np.einsum with widening) on a bounded row sample:
        # double the rows until one call takes a quarter of the CPU budget.
        sb = 0
        while True:
            sub = 1 << sb
            t0 = time.perf_counter()
            F.contract2(Lh[:sub].reshape((2,) * (sb + kb)), [100 + i for i in range(sb)] + K,
                        Rh.reshape((2,) * (nb + kb)), rw, [100 + i for i in range(sb)] + [200 + i for i in range(nb)],
                        args.mode)
            dt = time.perf_counter() - t0
            if dt > args.cpu_seconds / 4 or sb >= mb:
                break
            sb += 1
        cpu = {"value": 2.0 * sub * 2.0 ** (nb + kb) / dt / 1e12, "unit": "TOP/s", "cores": 1, "kind": "port",
               "sample": f"{sub} of {1 << mb} rows through the oracle's contract2 (np.einsum with widening), "
                         f"{dt:.1f}s, single thread"}
    peak = 5000.0
    line = {
        "metric": "TOP/s dense Boolean composition step (int8 MFMA, exact counts)",
        "value": tops, "unit": "TOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "i8", "data": "synthetic",
        "config": {"workload": f"dense composition 2^{mb} x 2^{kb} by 2^{kb} x 2^{nb}, K innermost, "
                               f"density 2^-{kb / 2:g}, {args.mode}",
                   "kernel_ms": kernel_ms, "valu_ms": valu_ms, "bytes": nbytes,
                   "hbm_gbs": nbytes / (kernel_ms / 1000.0) / 1e9,
                   "tile_autotune": None if tune is None else {
                       "chosen": "x".join(map(str, DT.TILES[tune["tile"]])),
                       "ms": {"x".join(map(str, DT.TILES[t])): v for t, v in tune["ms"].items()}}},
        "roofline": {"bound": "mfma", "achieved": tops, "peak": peak, "unit": "TFLOP/s", "frac": tops / peak,
                     "traffic": MEASURED_TRAFFIC.get(("dense", mb, nb, kb, tile), (None, None))[0],
                     "traffic_source": MEASURED_TRAFFIC.get(("dense", mb, nb, kb, tile), (None, None))[1],
                     "kernel": "k_step_mfma<%d,%d>" % (DT.TILES[tile][0] // 32, DT.TILES[tile][1] // 32)},
        "cpu_baseline": cpu,
        "parity": {"sampled_rows": len(rows), "mismatches": mism, "valu_mismatches": valu_mism},
    }
    print(json.dumps(line), flush=True)
    return 0 if mism == 0 and valu_mism == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
