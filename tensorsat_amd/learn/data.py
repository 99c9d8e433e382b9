"""
Training and evaluation data for the learned planner.

Instances follow the learned-planner workload: k uniform in {3, 4, 5}, n uniform in
[150, 300], m = round(0.9 * alpha_k * n) clauses of k distinct variables with random signs.
The generator here is vectorised (one numpy call per instance, duplicate clauses kept), so
its stream differs from ``CNFInstance.random`` (whose per-clause calls follow the
reference's generator); the distribution is the same.

Labels come from the plain planner: the cheapest of ``trials`` min-fill / min-degree
elimination orders of the instance's clause network (variables are the wires, clauses the
slots), each variable labelled with its normalised position in that order (0 = first
eliminated, 1 = last).
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from .. import plan as _plan
from .. import program as _prog
from ..workloads import THRESHOLD

TRAIN_BASE = 1_000_000
EVAL_BASE = 2_000_000


@dataclass
class KSat:
    num_vars: int
    lits: np.ndarray  # int32 [m, k], literals +-(v+1)

    @property
    def k(self) -> int:
        return int(self.lits.shape[1])

    @property
    def clauses(self) -> list[tuple[int, ...]]:
        return [tuple(int(x) for x in row) for row in self.lits]


def ksat(seed: int, k: int | None = None, n: int | None = None, ratio: float = 0.9) -> KSat:
    """One instance of the learned-planner workload from ``seed`` (k, n drawn unless given)."""
    rng = np.random.default_rng(seed)
    kk = int(rng.integers(3, 6))
    nn = int(rng.integers(150, 301))
    k = kk if k is None else k
    n = nn if n is None else n
    m = int(round(ratio * THRESHOLD[k] * n))
    # k distinct variables per clause: redraw rows that repeat a variable.
    v = rng.integers(0, n, size=(m, k))
    while True:
        s = np.sort(v, axis=1)
        bad = np.any(s[:, 1:] == s[:, :-1], axis=1)
        if not bad.any():
            break
        v[bad] = rng.integers(0, n, size=(int(bad.sum()), k))
    sign = rng.integers(0, 2, size=(m, k))
    lits = np.where(sign == 0, v + 1, -(v + 1)).astype(np.int32)
    return KSat(n, lits)


def _net(inst: KSat):
    """Planner network of the clause encoding: slot per clause over its variables; a
    variable in no clause gets a unary slot so that every wire is held."""
    n = inst.num_vars
    m, k = inst.lits.shape
    flat = (np.abs(inst.lits) - 1).astype(np.int32).reshape(-1)
    used = np.zeros(n, bool)
    used[flat] = True
    extra = np.nonzero(~used)[0].astype(np.int32)
    offs = np.concatenate([np.arange(m + 1, dtype=np.int32) * k, m * k + 1 + np.arange(len(extra), dtype=np.int32)])
    flat = np.concatenate([flat, extra]).astype(np.int32)
    dims = np.full(n, 2, np.int32)
    outw = np.zeros(1, np.int32)
    arrays = (np.ascontiguousarray(offs), np.ascontiguousarray(flat), dims, outw)
    net = _plan._Net(m + len(extra), _plan._ptr(arrays[0]), _plan._ptr(arrays[1]), n, _plan._ptr(dims), 0,
                     _plan._ptr(outw))
    return net, arrays


def plan_order(inst: KSat, trials: int = 8, seed: int = 0) -> tuple[np.ndarray, _plan.Cost]:
    """The plain planner's best elimination order (all variables) and its tree cost."""
    L = _prog._lib()
    net, keep = _net(inst)
    order = np.zeros(inst.num_vars, np.int32)
    c = _plan._Cost()
    _plan._check(L.tsp_elim_order(ctypes.byref(net), trials, seed, _plan._ptr(order), ctypes.byref(c)),
                 "tsp_elim_order")
    return order, _plan.Cost._from(c)


def prior_order(inst: KSat, prior, lam: float = 0.0) -> tuple[np.ndarray, _plan.Cost]:
    """One adaptive min-fill pass steered by per-variable ``prior`` (learned scores)."""
    L = _prog._lib()
    net, keep = _net(inst)
    pr = np.ascontiguousarray(np.asarray(prior, np.float32))
    if pr.shape != (inst.num_vars,):
        raise ValueError("one prior value per variable")
    order = np.zeros(inst.num_vars, np.int32)
    c = _plan._Cost()
    _plan._check(L.tsp_elim_order_prior(ctypes.byref(net), pr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                        float(lam), _plan._ptr(order), ctypes.byref(c)), "tsp_elim_order_prior")
    return order, _plan.Cost._from(c)


def order_cost(inst: KSat, order) -> _plan.Cost:
    """Exact tree cost of eliminating the variables in ``order``."""
    L = _prog._lib()
    net, keep = _net(inst)
    order = np.ascontiguousarray(np.asarray(order, np.int32))
    if order.shape != (inst.num_vars,):
        raise ValueError("order must list every variable")
    path = np.zeros(max(1, 2 * net.num_slots), np.int32)
    c = _plan._Cost()
    _plan._check(L.tsp_order_cost(ctypes.byref(net), _plan._ptr(order), _plan._ptr(path), ctypes.byref(c)),
                 "tsp_order_cost")
    return _plan.Cost._from(c)


def positions(order: np.ndarray) -> np.ndarray:
    n = len(order)
    pos = np.empty(n, np.float32)
    pos[order] = np.arange(n, dtype=np.float32) / max(n - 1, 1)
    return pos


def _label_one(args):
    seed, trials = args
    inst = ksat(seed)
    order, cost = plan_order(inst, trials=trials, seed=seed)
    return seed, positions(order), cost.log2_flops, cost.log2_max_size


def make_labels(base: int, count: int, trials: int = 8, procs: int | None = None, log=None) -> dict:
    """Labels for seeds base .. base+count-1 (a process pool over the host cores)."""
    import multiprocessing as mp

    procs = procs or min(16, os.cpu_count() or 1)
    seeds = list(range(base, base + count))
    out = {}
    with mp.get_context("fork").Pool(procs) as pool:
        for i, (s, pos, lf, lm) in enumerate(pool.imap_unordered(_label_one, [(s, trials) for s in seeds],
                                                                 chunksize=8)):
            out[s] = (pos, lf, lm)
            if log and (i + 1) % 1000 == 0:
                log(f"labels {i + 1}/{count}")
    nvar = np.asarray([len(out[s][0]) for s in seeds], np.int32)
    off = np.zeros(count + 1, np.int64)
    off[1:] = np.cumsum(nvar)
    return {
        "seeds": np.asarray(seeds, np.int64),
        "offsets": off,
        "labels": np.concatenate([out[s][0] for s in seeds]).astype(np.float16),
        "log2_flops": np.asarray([out[s][1] for s in seeds], np.float32),
        "log2_max": np.asarray([out[s][2] for s in seeds], np.float32),
        "trials": np.int64(trials),
    }


if __name__ == "__main__":
    import argparse
    import time

    ap = argparse.ArgumentParser(description="planner labels for the learned planner")
    ap.add_argument("--base", type=int, default=TRAIN_BASE)
    ap.add_argument("--count", type=int, default=20000)
    ap.add_argument("--trials", type=int, default=8)
    ap.add_argument("--procs", type=int, default=0)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    t = time.time()
    d = make_labels(a.base, a.count, a.trials, a.procs or None, log=lambda s: print(s, flush=True))
    np.savez(a.out, **d)
    print(f"{a.count} instances labelled in {time.time() - t:.1f}s -> {a.out}", flush=True)
