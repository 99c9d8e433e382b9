"""
Flat tensor networks (wirings) and the two SAT encodings.

A :class:`Wiring` mirrors the reference's flat ``Wiring`` (``tensorsat/diagrams.py:152-477``):
wire types (tensor extents), one wire list per slot, and the out wires.  Wires are
hyperedges: a wire on several slot ports is shared by all of them and, when it is
not an out wire, summed over (existentially quantified for Boolean relations).

Two encodings of a CNF instance are built here, both as flat index lists:

* ``bintree`` follows ``CNFInstance._build_diagram_bintree`` (``lib/sat.py:248-262``)
  composed after ``CNFInstance.inputs`` (``lib/sat.py:165-206``) and flattened
  (``diagrams.py:1452-1481``): slots are the input leaves (one per variable, in
  variable order), then per clause the ``not_`` gates of its negative literals,
  a balanced tree of ``or_`` gates and a ``bit_1`` assertion.  Wires ``0..n-1`` are
  the variables; every ``not_``/``or_`` gate adds one wire, in creation order.  This
  is the wire and slot numbering of ``(cnf.inputs(x) >> cnf.diagram()).flatten()``.
* ``clause`` puts one ``(2,)*k`` clause tensor per clause (a single 0 at the falsifying
  assignment) on the clause's distinct variables, with each variable a hyperedge
  (the copy/spider tensor of degree = occurrence count, kept implicit).  Contracting
  one clause's bintree sub-network gives exactly this clause tensor, so both
  encodings have the same verdict and the same model count.
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass, field

import numpy as np

from . import gates
from .sat import CNFInstance


class Wiring:
    """
    Immutable flat wiring (``diagrams.py:152-477``).

    :param wire_types: extent of each wire (``FinSet`` sizes; 2 for bits).
    :param slot_wires_list: for each slot, the wire on each of its ports.
    :param out_wires: the wire on each outer port (may repeat).
    """

    __slots__ = ("_wire_dims", "_slot_wires_list", "_out_wires", "_hash")

    def __init__(
        self,
        wire_types: Sequence[int],
        slot_wires_list: Sequence[Sequence[int]],
        out_wires: Sequence[int],
    ) -> None:
        wire_dims = tuple(int(d) for d in wire_types)
        if any(d <= 0 for d in wire_dims):
            raise ValueError("Wire extents must be strictly positive.")
        slot_wires_list = tuple(tuple(int(w) for w in ws) for ws in slot_wires_list)
        out_wires = tuple(int(w) for w in out_wires)
        nw = len(wire_dims)
        for slot, ws in enumerate(slot_wires_list):
            for w in ws:
                if not 0 <= w < nw:
                    raise ValueError(f"Invalid wire index {w} in slot mapping for slot {slot}.")
        for w in out_wires:
            if not 0 <= w < nw:
                raise ValueError(f"Invalid wire index {w} in outer mapping.")
        self._wire_dims = wire_dims
        self._slot_wires_list = slot_wires_list
        self._out_wires = out_wires
        self._hash = hash((wire_dims, slot_wires_list, out_wires))

    @property
    def wire_types(self) -> tuple[int, ...]:
        return self._wire_dims

    wire_dims = wire_types

    @property
    def slot_wires_list(self) -> tuple[tuple[int, ...], ...]:
        return self._slot_wires_list

    @property
    def out_wires(self) -> tuple[int, ...]:
        return self._out_wires

    @property
    def num_wires(self) -> int:
        return len(self._wire_dims)

    @property
    def num_slots(self) -> int:
        return len(self._slot_wires_list)

    @property
    def slot_shapes(self) -> tuple[tuple[int, ...], ...]:
        d = self._wire_dims
        return tuple(tuple(d[w] for w in ws) for ws in self._slot_wires_list)

    @property
    def shape(self) -> tuple[int, ...]:
        return tuple(self._wire_dims[w] for w in self._out_wires)

    @property
    def dangling_wires(self) -> frozenset[int]:
        """Wires on no slot port (``diagrams.py:290-295``)."""
        used = {w for ws in self._slot_wires_list for w in ws}
        return frozenset(range(self.num_wires)) - used

    @property
    def discarded_wires(self) -> frozenset[int]:
        return frozenset(range(self.num_wires)) - set(self._out_wires)

    def wired_slots(self) -> dict[int, tuple[int, ...]]:
        res: dict[int, list[int]] = {}
        for s, ws in enumerate(self._slot_wires_list):
            for w in ws:
                res.setdefault(w, []).append(s)
        return {w: tuple(v) for w, v in res.items()}

    def flat_arrays(self) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        """(slot_offsets int32[S+1], slot_wire_ids int32[...], out_wires int32, wire_dims int32)."""
        offs = np.zeros(self.num_slots + 1, dtype=np.int32)
        offs[1:] = np.cumsum([len(ws) for ws in self._slot_wires_list])
        flat = np.fromiter(
            (w for ws in self._slot_wires_list for w in ws), dtype=np.int32, count=int(offs[-1])
        )
        return (
            offs,
            flat,
            np.asarray(self._out_wires, dtype=np.int32),
            np.asarray(self._wire_dims, dtype=np.int32),
        )

    def __eq__(self, other: object) -> bool:
        if not isinstance(other, Wiring):
            return NotImplemented
        return (
            self._hash == other._hash
            and self._wire_dims == other._wire_dims
            and self._slot_wires_list == other._slot_wires_list
            and self._out_wires == other._out_wires
        )

    def __hash__(self) -> int:
        return self._hash

    def __repr__(self) -> str:
        attrs = []
        if self.num_wires:
            attrs.append(f"{self.num_wires} wire{'s' if self.num_wires != 1 else ''}")
        if self.num_slots:
            attrs.append(f"{self.num_slots} slot{'s' if self.num_slots != 1 else ''}")
        if self._out_wires:
            n = len(self._out_wires)
            attrs.append(f"{n} out port{'s' if n != 1 else ''}")
        return f"<Wiring: {', '.join(attrs)}>" if attrs else "<Wiring (empty)>"


@dataclass(frozen=True)
class Diagram:
    """
    A flat diagram: a wiring plus one block per slot (``diagrams.py:1216``).

    Blocks are host uint8 arrays (leaf tensors) or boxes of any box class exposing
    ``tensor_shape``; contractions convert them to their own box class.
    """

    wiring: Wiring
    blocks: tuple = field(default=())
    leaf_kinds: tuple[str, ...] | None = None
    _cnf_source: tuple | None = field(default=None, compare=False, repr=False)

    def __post_init__(self) -> None:
        if len(self.blocks) != self.wiring.num_slots:
            raise ValueError("One block per slot is required.")
        for s, (b, shape) in enumerate(zip(self.blocks, self.wiring.slot_shapes)):
            bshape = tuple(getattr(b, "tensor_shape", None) or np.shape(b))
            if bshape != shape:
                raise ValueError(
                    f"Incompatible shape for data at slot {s}: expected slot shape"
                    f" {shape}, got data with shape {bshape}."
                )

    @property
    def boxes(self) -> tuple:
        return self.blocks

    @property
    def num_open_slots(self) -> int:
        return 0

    @property
    def is_flat(self) -> bool:
        return True

    def flatten(self) -> "Diagram":
        """Flat already (``diagrams.py:1452-1481`` is the identity on flat diagrams)."""
        return self

    def __repr__(self) -> str:
        w = self.wiring
        attrs = [f"{w.num_wires} wires", f"{w.num_slots} blocks"]
        if w.out_wires:
            attrs.append(f"{len(w.out_wires)} ports")
        return f"<Diagram: {', '.join(attrs)}>"


_INPUT_LEAF = {"0": ("bit_0", gates.bit_0), "1": ("bit_1", gates.bit_1), "?": ("bit_unk", gates.bit_unk)}


def _spec(cnf: CNFInstance, inputs) -> str:
    """Input selection: ``0``/``1``/``?`` leaves, ``_`` open, ``*`` summed without a leaf."""
    if isinstance(inputs, str) and "*" in inputs:
        if len(inputs) != cnf.num_vars or not all(b in "01?_*" for b in inputs):
            raise ValueError("Input specification characters must be '0', '1', '?', '_' or '*'.")
        return inputs
    return cnf.normalize_inputs(inputs)


def bintree_network(cnf: CNFInstance, inputs: str | int | None = None) -> Diagram:
    """
    ``(cnf.inputs(inputs) >> cnf.diagram()).flatten()`` as a flat diagram
    (``lib/sat.py:165-262``; composition ``diagrams.py:1592-1637``).

    ``'_'`` leaves a variable open: no input leaf, and its wire becomes an out wire,
    as in contracting ``cnf.diagram()`` itself (its ports are the variables); ``'*'``
    sums it without a leaf (``inputs(..., discard_unk=True)``).
    """
    spec = _spec(cnf, inputs)
    n = cnf.num_vars
    slot_wires: list[tuple[int, ...]] = []
    blocks: list[np.ndarray] = []
    kinds: list[str] = []
    out_wires: list[int] = []
    for i, b in enumerate(spec):
        if b == "_":
            out_wires.append(i)
        elif b != "*":
            kind, leaf = _INPUT_LEAF[b]
            slot_wires.append((i,))
            blocks.append(leaf)
            kinds.append(kind)
    next_wire = n
    for clause in cnf.clauses:
        layer: list[int] = []
        for lit in clause:
            if lit > 0:
                layer.append(lit - 1)
            else:
                slot_wires.append((-lit - 1, next_wire))
                blocks.append(gates.not_)
                kinds.append("not_")
                layer.append(next_wire)
                next_wire += 1
        while (ln := len(layer)) > 1:
            new_layer = []
            for j in range(ln // 2):
                slot_wires.append((layer[2 * j], layer[2 * j + 1], next_wire))
                blocks.append(gates.or_)
                kinds.append("or_")
                new_layer.append(next_wire)
                next_wire += 1
            if ln % 2 == 1:
                slot_wires.append((new_layer[-1], layer[-1], next_wire))
                blocks.append(gates.or_)
                kinds.append("or_")
                new_layer[-1] = next_wire
                next_wire += 1
            layer = new_layer
        if len(layer) != 1:
            raise ValueError("Empty clauses have no bintree encoding.")
        slot_wires.append((layer[0],))
        blocks.append(gates.bit_1)
        kinds.append("bit_1")
    wiring = Wiring((2,) * next_wire, slot_wires, out_wires)
    return Diagram(wiring, tuple(blocks), tuple(kinds))


def clause_network(cnf: CNFInstance, inputs: str | int | None = None) -> Diagram:
    """
    Clause encoding: one clause tensor per clause over its distinct variables, with
    variables as hyperedges (wire ``i`` = variable ``i``).

    * A clause holding both ``x`` and ``-x`` is always true: it contributes an all-ones
      tensor on its distinct variables (kept so that no variable loses its wire).
    * An empty clause is the 0-d tensor ``0``.
    * A variable in no clause gets a ``bit_unk`` leaf under ``'?'`` so that model
      counts include its factor 2.
    * ``'0'``/``'1'`` add a ``bit_0``/``bit_1`` leaf on the variable; ``'_'`` makes it an
      out wire.
    """
    spec = _spec(cnf, inputs)
    n = cnf.num_vars
    slot_wires: list[tuple[int, ...]] = []
    blocks: list[np.ndarray] = []
    kinds: list[str] = []
    used = np.zeros(n, dtype=bool)
    cache: dict[tuple[bool, ...], np.ndarray] = {}
    for clause in cnf.clauses:
        pol: dict[int, set[bool]] = {}
        order: list[int] = []
        for lit in clause:
            v = abs(lit) - 1
            if v not in pol:
                pol[v] = set()
                order.append(v)
            pol[v].add(lit > 0)
        if any(len(p) == 2 for p in pol.values()):
            t = np.ones((2,) * len(order), dtype=np.uint8)
            kind = "taut"
        else:
            signs = tuple(next(iter(pol[v])) for v in order)
            if signs not in cache:
                ct = gates.clause_tensor(signs)
                ct.setflags(write=False)
                cache[signs] = ct
            t = cache[signs]
            kind = "clause"
        slot_wires.append(tuple(order))
        blocks.append(t)
        kinds.append(kind)
        used[order] = True
    out_wires: list[int] = []
    for i, b in enumerate(spec):
        if b == "_":
            out_wires.append(i)
        elif b == "*" and used[i]:
            continue
        elif b in "01" or not used[i]:
            kind, leaf = _INPUT_LEAF["?" if b == "*" else b]
            slot_wires.append((i,))
            blocks.append(leaf)
            kinds.append(kind)
    wiring = Wiring((2,) * n, slot_wires, out_wires)
    return Diagram(wiring, tuple(blocks), tuple(kinds))


def sat_network(cnf: CNFInstance, inputs: str | int | None = None, *, mode: str = "bintree") -> Diagram:
    """Dispatch on the encoding name: ``"bintree"`` (reference) or ``"clause"``."""
    if mode == "bintree":
        return bintree_network(cnf, inputs)
    if mode == "clause":
        return clause_network(cnf, inputs)
    raise NotImplementedError(f"Unknown diagram mode for CNFInstance: {mode!r}")
