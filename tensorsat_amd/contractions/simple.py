"""
Explicit-path contractions (``tensorsat/contractions/simple.py:75-294``).

The contract2 argument lists are precomputed exactly as the reference's
``SimpleContraction._new`` does (``simple.py:134-194``): at each step the result keeps
``lhs_only + common + rhs_only`` wires, and every wire that is not an out wire is dropped
from the result of the last step in which it appears.  ``contract`` then runs those
steps, adds dangling out wires with ``spider(t, 1)`` products and rewires to the out
wires (``simple.py:251-292``).  Paths use the opt_einsum convention (pop two, append).
"""

from __future__ import annotations

from collections.abc import Sequence

from .. import plan as _plan
from ..network import Diagram, Wiring
from .abc import Contraction

ContractionPath = tuple[tuple[int, int], ...]


class SimpleContraction(Contraction):
    """A contraction along an explicit pop/append path."""

    def __init__(self, box_class: type, wiring: Wiring, path: Sequence[tuple[int, int]]) -> None:
        path = tuple((int(a), int(b)) for a, b in path)
        super().__init__(box_class, wiring)
        if wiring.num_wires == 0:
            raise ValueError("Cannot define contraction for empty wiring.")
        n = wiring.num_slots
        if not path and n >= 2:
            raise ValueError("Contraction path can only be empty for number of wiring slots <= 1.")
        if len(path) >= n and n > 0:
            raise ValueError(
                "Path is too long for the given wiring:"
                f" there are {len(path)} contractions, but {n} slots to contract."
            )
        if n > 0 and len(path) != n - 1:
            raise ValueError(
                "Path does not fully contract the given wiring:"
                f" there are {len(path)} contractions, but {n} slots to contract."
            )
        for idx, (lhs, rhs) in enumerate(path):
            if not 0 <= lhs < n - idx:
                raise ValueError(f"Invalid lhs for contraction {(lhs, rhs) = } at {idx = }.")
            if not 0 <= rhs < n - idx:
                raise ValueError(f"Invalid rhs for contraction {(lhs, rhs) = } at {idx = }.")
        self._path = path
        self._compile()

    @classmethod
    def using_planner(cls, box_class: type, wiring: Wiring, optimize: str = "greedy") -> "SimpleContraction":
        """A simple contraction along a path from this package's planner (the role of
        ``using_opt_einsum``, ``simple.py:79-116``; opt_einsum is not a dependency)."""
        from .planned import plan_wiring

        return cls(box_class, wiring, plan_wiring(wiring, optimize).linear_path())

    using_opt_einsum = using_planner

    @classmethod
    def contract_using_planner(cls, box_class: type, diagram: Diagram, optimize: str = "greedy"):
        return cls.using_planner(box_class, diagram.wiring, optimize).contract(diagram)

    contract_using_opt_einsum = contract_using_planner

    def _compile(self) -> None:
        wiring = self.wiring
        path = self._path
        if path:
            args: list[list] = []
            wires_list = [tuple(ws) for ws in wiring.slot_wires_list]
            for lhs_idx, rhs_idx in path:
                lhs_wires = wires_list.pop(lhs_idx)
                rhs_wires = wires_list.pop(rhs_idx - int(lhs_idx < rhs_idx))
                common = set(lhs_wires) & set(rhs_wires)
                if common:
                    lhs_only = tuple(w for w in lhs_wires if w not in common)
                    rhs_only = tuple(w for w in rhs_wires if w not in common)
                    common_w = tuple(w for w in lhs_wires if w in common)
                    res_wires = lhs_only + common_w + rhs_only
                else:
                    res_wires = lhs_wires + rhs_wires
                wires_list.append(res_wires)
                args.append([lhs_idx, lhs_wires, rhs_idx, rhs_wires, list(res_wires)])
            final_res_wires = list(args[-1][-1])
            out_wires = wiring.out_wires
            for w in sorted(set(final_res_wires) - set(out_wires)):
                for entry in reversed(args):
                    if w in entry[-1]:
                        entry[-1].remove(w)
                        break
            self._contract2_args = tuple(
                (a, tuple(lw), b, tuple(rw), tuple(res)) for a, lw, b, rw, res in args
            )
            self._box_out_wires = tuple(args[-1][-1])
        else:
            self._contract2_args = ()
            if wiring.num_slots == 0:
                self._box_out_wires = ()
            else:
                self._box_out_wires = tuple(wiring.slot_wires_list[0])
        self._dangling = tuple(sorted(wiring.dangling_wires))

    @property
    def path(self) -> ContractionPath:
        return self._path

    @property
    def contract2_args(self):
        return self._contract2_args

    def contract(self, diagram: Diagram):
        """``simple.py:251-292``."""
        self.validate(diagram)
        cls = self.box_class
        wiring = self.wiring
        boxes = self._boxes(diagram)
        box = None
        box_out_wires = self._box_out_wires
        if boxes:
            for lhs_idx, lhs_wires, rhs_idx, rhs_wires, res_wires in self._contract2_args:
                lhs = boxes.pop(lhs_idx)
                rhs = boxes.pop(rhs_idx - int(lhs_idx < rhs_idx))
                boxes.append(cls._contract2(lhs, lhs_wires, rhs, rhs_wires, res_wires))
            box = boxes[0]
        if self._dangling:
            box_out_wires = box_out_wires + self._dangling
            parts = ([box] if box is not None else []) + [
                cls.spider(wiring.wire_dims[w], 1) for w in self._dangling
            ]
            box = cls.prod(parts)
        if box is None:
            box = cls.scalar(True)
        out_wires = wiring.out_wires
        if tuple(out_wires) != tuple(box_out_wires):
            box = box._rewire([box_out_wires.index(w) for w in out_wires])
        return box
