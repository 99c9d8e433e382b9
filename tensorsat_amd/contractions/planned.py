"""
Planned (tree) contractions: the counterpart of the reference's ``CotengraContraction``
(``tensorsat/contractions/cotengra.py:52-131``) with this package's planner in place of
cotengra.

Construction plans a binary tree over the wiring's hypergraph (``optimize``: ``"greedy"``,
``"hyper"``, ``"elim"`` or ``"auto"``, the cheapest of elimination-order and hyper-greedy
trees by the cotengra cost model), optionally slicing wires down to a size target.
``contract`` walks the tree calling the box class's pairwise contraction per node, as
cotengra's ``ContractionTree.contract`` does through ``_autoray.einsum`` (a wire is
summed at the first node after which no other tensor and no out port holds it), OR-s
(``+``) the slices, adds dangling out wires and rewires to the out wires.
:meth:`contract_program` compiles the whole tree for the device executor instead.
"""

from __future__ import annotations

import itertools

import numpy as np

from .. import plan as _plan
from .. import program as _prog
from ..network import Diagram, Wiring
from .abc import Contraction


def plan_wiring(wiring: Wiring, optimize: str = "auto", *, trials: int = 32, seed: int = 0) -> _plan.Plan:
    if optimize == "greedy":
        return _plan.greedy(wiring)
    if optimize == "hyper":
        return _plan.hyper(wiring, trials=trials, seed=seed)
    if optimize == "elim":
        p = _prog.elim(wiring, trials=max(1, trials // 4), seed=seed)
        p.cost = _plan.tree_cost(wiring, p)
        return p
    if optimize in ("auto", "auto-hq"):
        cands = [_plan_elim(wiring, trials, seed), _plan.hyper(wiring, trials=trials, seed=seed)]
        for c in cands:
            c.cost = _plan.tree_cost(wiring, c)
        return min(cands, key=lambda p: (p.cost.log2_max_size, p.cost.log2_flops))
    raise ValueError(f"unknown optimize {optimize!r}")


def _plan_elim(wiring: Wiring, trials: int, seed: int) -> _plan.Plan:
    return _prog.elim(wiring, trials=max(1, trials // 4), seed=seed)


def tree_steps(wiring: Wiring, path: np.ndarray, drop=()):
    """
    Per merge: (lhs id, rhs id, lhs wires, rhs wires, out wires) under the merge rule,
    with wires in ``drop`` (sliced) removed everywhere.  Out wires keep lhs order then
    rhs-only order.
    """
    drop = set(drop)
    S = wiring.num_slots
    legs: list[list[int]] = []
    count: dict[int, int] = {}
    for ws in wiring.slot_wires_list:
        u = []
        for w in ws:
            if w not in drop and w not in u:
                u.append(w)
        legs.append(u)
        for w in u:
            count[w] = count.get(w, 0) + 1
    for w in set(wiring.out_wires):
        count[w] = count.get(w, 0) + 1
    steps = []
    for a, b in path.tolist():
        la, lb = legs[a], legs[b]
        sb = set(lb)
        out = []
        for w in la + [w for w in lb if w not in set(la)]:
            c = (w in set(la)) + (w in sb)
            if count.get(w, 0) > c:
                out.append(w)
        for w in la:
            count[w] -= 1
        for w in lb:
            count[w] -= 1
        for w in out:
            count[w] = count.get(w, 0) + 1
        legs.append(out)
        steps.append((a, b, la, lb, out))
    root_legs = legs[-1] if legs else []
    return steps, root_legs


class PlannedContraction(Contraction):
    """A contraction along a planned binary tree (``CotengraContraction`` counterpart)."""

    def __init__(self, box_class: type, wiring: Wiring, *, optimize: str = "auto",
                 slicing_target: float | None = None, trials: int = 32, seed: int = 0,
                 canonicalize: bool = True, sort_contraction_indices: bool = False) -> None:
        super().__init__(box_class, wiring)
        if wiring.num_wires == 0:
            raise ValueError("Cannot define contraction for empty wiring.")
        tree = plan_wiring(wiring, optimize, trials=trials, seed=seed)
        if slicing_target is not None and tree.cost is not None and tree.cost.log2_max_size > slicing_target:
            tree = _plan.slice_plan(wiring, tree, slicing_target, reconfigure_trials=trials, seed=seed)
        self._tree = tree

    @property
    def tree(self) -> _plan.Plan:
        return self._tree

    @property
    def path(self) -> tuple[tuple[int, int], ...]:
        return tuple(self._tree.linear_path())

    def _contract_boxes(self, boxes: list, fixed: dict[int, int]):
        cls = self.box_class
        wiring = self.wiring
        leaves = []
        for b, ws in zip(boxes, wiring.slot_wires_list):
            if fixed and any(w in fixed for w in ws):
                idx = tuple(fixed[w] if w in fixed else slice(None) for w in ws)
                # a sliced wire repeated within one slot is fixed on each of its ports
                b = b[idx]
            leaves.append(b)
        live = {}
        lw = {}
        for s, (b, ws) in enumerate(zip(leaves, wiring.slot_wires_list)):
            ws = [w for w in ws if w not in fixed]
            # repeated wires within one slot: take the diagonal first
            uniq = list(dict.fromkeys(ws))
            if len(uniq) != len(ws):
                b = cls._contract2(b, ws, cls.scalar(True), [], uniq)
            live[s] = b
            lw[s] = uniq
        steps, root = tree_steps(wiring, self._tree.path, fixed)
        S = wiring.num_slots
        for k, (a, b, la, lb, out) in enumerate(steps):
            res = cls._contract2(live.pop(a), lw.pop(a), live.pop(b), lw.pop(b), out)
            live[S + k] = res
            lw[S + k] = out
        if S == 0:
            box, wires = cls.scalar(True), []
        else:
            (key,) = live.keys()
            box, wires = live[key], list(lw[key])
            # a single slot: sum the wires that are not out wires
            keep = [w for w in wires if w in set(wiring.out_wires)]
            if keep != wires:
                box = box._rewire([wires.index(w) for w in keep])
                wires = keep
        dangling = [w for w in dict.fromkeys(wiring.out_wires) if w not in wires and w not in fixed]
        if dangling:
            box = cls.prod([box] + [cls.spider(wiring.wire_dims[w], 1) for w in dangling])
            wires = wires + dangling
        return box, wires

    def contract(self, diagram: Diagram, *, order=None, autojit="auto", progbar: bool = False):
        """Validate and contract (``cotengra.py:100-128``)."""
        self.validate(diagram)
        boxes = self._boxes(diagram)
        sliced = list(self._tree.sliced)
        dims = self.wiring.wire_dims
        result = None
        wires_out = None
        for vals in itertools.product(*(range(dims[w]) for w in sliced)):
            box, wires = self._contract_boxes(boxes, dict(zip(sliced, vals)))
            result = box if result is None else result + box
            wires_out = wires
        out_wires = list(self.wiring.out_wires)
        if wires_out != out_wires:
            result = result._rewire([wires_out.index(w) for w in out_wires])
        return result

    def contract_program(self, diagram: Diagram) -> _prog.Program:
        """
        Compile the tree with the diagram's host leaves for the device step executors:
        one job, or with sliced wires one job per slice (the sliced wires fixed in every
        leaf that holds them, ``Box.__getitem__``; the jobs' results are to be OR-ed, or
        summed when counting).  Every job has the tree's steps with the sliced wires
        removed, in the compiler's layouts (contracted wires innermost in both operands,
        so dense steps take the MFMA route).  Needs extent-2 wires, unique wires per slot
        and a scalar result.
        """
        self.validate(diagram)
        if self.wiring.out_wires:
            raise NotImplementedError("compiled programs of sliced trees need a scalar result")
        leaves = [np.asarray(getattr(b, "tensor", b), dtype=np.uint8) for b in diagram.blocks]
        sliced = list(self._tree.sliced)
        if not sliced:
            return _prog.compile_network(self.wiring, leaves, self._tree)
        progs = _prog.compile_sliced(self.wiring, leaves, self._tree, sliced)
        return _prog.concat(progs, [np.zeros(1, np.int64)] * len(progs), 1)

    def __repr__(self) -> str:
        return "<PlannedContraction>"
