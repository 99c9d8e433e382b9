"""
The learned planner at serving time: one batched GPU inference of the trained network
gives every variable a score, which the cube search uses as its proposal.

* Trees: each cube's residual network is planned by one min-fill pass steered by the
  scores (``tsp_elim_order_prior``); the exact cost model keeps the proposal only if it is
  no larger than one deterministic min-fill pass and no costlier by more than the gate
  margin, otherwise the plain multi-trial planner is the fallback.
* Slices: with ``branch=True`` the search conditions on the free variable of highest
  score (the variables the planner eliminates last, i.e. those in the widest bags).

The proposal never changes an answer: every cube is still contracted exactly and the
verdict is the OR over cubes, whichever tree or branching order produced them.
"""

from __future__ import annotations

import time
from pathlib import Path

import numpy as np

from . import graph as G

WEIGHTS = Path(__file__).resolve().parent / "weights" / "ksat_gnn.npz"


class LearnedPlanner:
    def __init__(self, weights: str | Path = WEIGHTS, device: str = "cuda") -> None:
        from .model import GNN

        self.model = GNN.load(weights, device=device)
        self.last_infer_s = 0.0

    def scores(self, cnfs) -> list[np.ndarray]:
        """Per-instance score vectors (instances with ``num_vars`` and ``clauses`` or
        ``lits``), from one batched inference."""
        import torch

        t = time.perf_counter()
        g = G.build(cnfs)
        dg = G.DeviceGraph(g)
        y, _ = self.model.forward(dg)
        y = y.cpu().numpy()
        torch.cuda.synchronize()
        self.last_infer_s = time.perf_counter() - t
        return [y[g.var_offset[i] : g.var_offset[i + 1]].copy() for i in range(len(cnfs))]

    def priority(self, cnfs) -> np.ndarray:
        """Scores of all instances concatenated (the ``priority`` argument of
        ``program.sat_batch``)."""
        s = self.scores(cnfs)
        return np.concatenate(s).astype(np.float32) if s else np.zeros(0, np.float32)
