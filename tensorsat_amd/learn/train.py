"""
Training of the learned planner on the GPU, with the project's own forward and backward
kernels (``ts_gnn_*``), and evaluation of its proposals with the exact cost model.

Batches of instances are generated from their seeds and turned into incidence graphs by a
host process pool while the GPU trains; the target of each variable is its normalised
position in the plain planner's best elimination order (``data.make_labels``), the loss
the mean squared error, the optimiser Adam with a cosine learning-rate decay.

Evaluation on held-out seeds: the proposal is the order of ascending score; its exact tree
cost (largest intermediate, flops) is compared with the planner's best-of-``trials`` order
(the labels) and with one deterministic min-fill pass, and the time to produce each is
recorded (one batched GPU inference for all instances against the host planner).
"""

from __future__ import annotations

import math
import os
import time

import numpy as np

from . import data as D
from . import graph as G


def _batch_graph(seeds):
    return G.build([D.ksat(int(s)) for s in seeds])


def _pool(procs: int):
    """Host worker pool; spawned, so no worker inherits the parent's GPU state."""
    import multiprocessing as mp

    return mp.get_context("spawn").Pool(procs)


def load_labels(path):
    with np.load(path, allow_pickle=False) as z:
        return z["seeds"].copy(), z["offsets"].copy(), z["labels"].astype(np.float32), \
            z["log2_flops"].copy(), z["log2_max"].copy()


def train(labels_path, *, epochs: int = 3, batch: int = 32, lr: float = 1e-3, procs: int = 12, seed: int = 0,
          layers: int | None = None, log=print, max_seconds: float | None = None):
    import torch

    from .model import GNN, LAYERS

    seeds, off, lab, _, _ = load_labels(labels_path)
    model = GNN(seed=seed, layers=layers or LAYERS)
    rng = np.random.default_rng(seed)
    steps_per_epoch = (len(seeds) + batch - 1) // batch
    total = epochs * steps_per_epoch
    step = 0
    t0 = time.perf_counter()
    history = []
    loss_acc = torch.zeros(1, dtype=torch.float32, device="cuda")
    n_acc = 0
    done = False
    with _pool(procs) as pool:
        for ep in range(epochs):
            perm = rng.permutation(len(seeds))
            groups = [perm[i : i + batch] for i in range(0, len(perm), batch)]
            for gi, g in enumerate(pool.imap(_batch_graph, [seeds[b] for b in groups], chunksize=1)):
                b = groups[gi]
                target = np.concatenate([lab[off[i] : off[i + 1]] for i in b])
                dg = G.DeviceGraph(g)
                t = torch.from_numpy(target).cuda()
                _, loss = model.forward(dg, target=t, train=True)
                cur = lr * (0.1 + 0.9 * 0.5 * (1.0 + math.cos(math.pi * step / max(total, 1))))
                model.adam_step(lr=cur)
                loss_acc += loss
                n_acc += 1
                step += 1
                if step % 50 == 0 or step == total:
                    mean = float(loss_acc.item()) / n_acc
                    history.append((step, mean))
                    log(f"epoch {ep} step {step}/{total} loss {mean:.5f} lr {cur:.2e} "
                        f"{time.perf_counter() - t0:.1f}s")
                    loss_acc.zero_()
                    n_acc = 0
                if max_seconds and time.perf_counter() - t0 > max_seconds:
                    log(f"time limit reached at step {step}")
                    done = True
                    break
            if done:
                break
    torch.cuda.synchronize()
    return model, {"steps": step, "seconds": time.perf_counter() - t0, "history": history}


def scores(model, instances, batch: int = 64):
    """Per-instance score vectors (one batched inference per ``batch`` instances)."""
    import torch

    out = []
    for i in range(0, len(instances), batch):
        part = instances[i : i + batch]
        g = G.build(part)
        dg = G.DeviceGraph(g)
        y, _ = model.forward(dg)
        y = y.cpu().numpy()
        for j in range(len(part)):
            out.append(y[g.var_offset[j] : g.var_offset[j + 1]].copy())
    torch.cuda.synchronize()
    return out


def _spearman(a, b):
    ra = np.argsort(np.argsort(a)).astype(np.float64)
    rb = np.argsort(np.argsort(b)).astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra * rb).sum() / max(np.sqrt((ra * ra).sum() * (rb * rb).sum()), 1e-12))


def _plain_one(seed):
    inst = D.ksat(int(seed))
    t = time.perf_counter()
    _, c = D.plan_order(inst, trials=1, seed=0)
    return c.log2_max_size, c.log2_flops, time.perf_counter() - t


def _order_one(args):
    seed, sc = args
    inst = D.ksat(int(seed))
    c = D.order_cost(inst, np.argsort(sc, kind="stable").astype(np.int32))
    return c.log2_max_size, c.log2_flops


def evaluate(model, labels_path, *, procs: int = 12, log=print) -> dict:
    """Plan quality of the proposals on held-out seeds (labels file of those seeds)."""
    import torch

    seeds, off, lab, lab_flops, lab_max = load_labels(labels_path)
    insts = [D.ksat(int(s)) for s in seeds]
    torch.cuda.synchronize()
    t = time.perf_counter()
    sc = scores(model, insts)
    infer_s = time.perf_counter() - t
    rho = [_spearman(sc[i], lab[off[i] : off[i + 1]]) for i in range(len(seeds))]
    with _pool(procs) as pool:
        learned = pool.map(_order_one, [(int(s), sc[i]) for i, s in enumerate(seeds)], chunksize=4)
        plain = pool.map(_plain_one, [int(s) for s in seeds], chunksize=4)
    lm = np.array([x[0] for x in learned])
    lf = np.array([x[1] for x in learned])
    pm = np.array([x[0] for x in plain])
    pf = np.array([x[1] for x in plain])
    plain_s = float(sum(x[2] for x in plain))
    ks = np.array([x.k for x in insts])
    res = {
        "instances": len(seeds),
        "spearman_mean": float(np.mean(rho)),
        "learned_log2_max_mean": float(lm.mean()),
        "planner_best_log2_max_mean": float(lab_max.mean()),
        "minfill_log2_max_mean": float(pm.mean()),
        "learned_log2_flops_mean": float(lf.mean()),
        "planner_best_log2_flops_mean": float(lab_flops.mean()),
        "minfill_log2_flops_mean": float(pf.mean()),
        "learned_le_minfill_frac": float(np.mean(lm <= pm)),
        "inference_s": infer_s,
        "minfill_host_s": plain_s,
        "per_k": {int(k): {"learned_log2_max": float(lm[ks == k].mean()), "planner_log2_max": float(lab_max[ks == k].mean()),
                           "minfill_log2_max": float(pm[ks == k].mean())} for k in sorted(set(ks.tolist()))},
    }
    log(f"eval: spearman {res['spearman_mean']:.3f}; log2 max learned {res['learned_log2_max_mean']:.1f} "
        f"planner {res['planner_best_log2_max_mean']:.1f} min-fill {res['minfill_log2_max_mean']:.1f}; "
        f"inference {infer_s:.2f}s for {len(seeds)} vs min-fill {plain_s:.2f}s host")
    return res


if __name__ == "__main__":
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description="train the learned planner on the GPU")
    ap.add_argument("--labels", required=True, help="labels file (data.py) of the training seeds")
    ap.add_argument("--eval-labels", help="labels file of held-out seeds")
    ap.add_argument("--out", required=True, help="weights file to write (.npz)")
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--procs", type=int, default=min(12, os.cpu_count() or 1))
    ap.add_argument("--max-seconds", type=float, default=None)
    ap.add_argument("--report", help="write the training/evaluation summary (JSON) here")
    a = ap.parse_args()
    say = lambda s: print(s, flush=True)
    model, info = train(a.labels, epochs=a.epochs, batch=a.batch, lr=a.lr, procs=a.procs, log=say,
                        max_seconds=a.max_seconds)
    model.save(a.out)
    say(f"saved {a.out} after {info['steps']} steps in {info['seconds']:.1f}s")
    rep = {"train": info}
    if a.eval_labels:
        rep["eval"] = evaluate(model, a.eval_labels, procs=a.procs, log=say)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    sys.exit(0)
