"""
Binary circuits as flat tensor networks: the half and full adders, the ripple-carry
adder and the Wallace-tree multiplier of the reference's ``tensorsat/lib/bincirc.py:91-177``.

The reference builds these with ``DiagramBuilder`` recipes (``diagrams.py:1719-1771``)
whose adder blocks are sub-diagrams, and a contraction needs the flattened diagram
(``Diagram.flatten``, ``diagrams.py:1452-1481``, composing wirings by
``Wiring._compose``, ``diagrams.py:387-461``).  :class:`CircuitBuilder` records the same
construction -- inputs, blocks applied to wires (``block @ diag[wires]``: the selected
wires go to the block's first ports, every other port gets a new wire), outputs -- and
produces the flat :class:`~tensorsat_amd.network.Diagram` directly, numbering wires as
the reference's flattening does:

1. the builder's own wires keep their order, except that wires joined through a
   sub-circuit (one inner wire on several of its ports) share the number of the lowest
   of them;
2. then, sub-circuit by sub-circuit in slot order, each inner wire not on one of its
   ports gets the next number, in inner-wire order;
3. a sub-circuit's slot is replaced by its own slots (inner wires renumbered), a gate's
   slot keeps its place.

So ``rc_adder(n).diagram.wiring`` equals the reference's ``rc_adder(n).flatten().wiring``
and its ports are ``a[0..n-1], b[0..n-1], c_in, s[0..n-1], c_out``; contracting it with any
box class gives the adder's relation.  :meth:`Circuit.summary` reports the nested
diagram's wire, block and port counts (the reference's ``repr`` of the recipe result).
"""

from __future__ import annotations

from collections.abc import Iterable, Sequence
from dataclasses import dataclass

import numpy as np

from . import gates
from .network import Diagram, Wiring


@dataclass(frozen=True)
class Circuit:
    """A built circuit: the flat diagram (gate tensors as blocks, their names as
    ``leaf_kinds``) and the nested construction's counts."""

    name: str
    diagram: Diagram
    nested_wires: int
    nested_blocks: int
    depth: int

    @property
    def wiring(self) -> Wiring:
        return self.diagram.wiring

    @property
    def num_ports(self) -> int:
        return len(self.diagram.wiring.out_wires)

    def summary(self) -> dict:
        """Counts of the nested diagram, as the reference's ``repr`` prints them."""
        return {"wires": self.nested_wires, "blocks": self.nested_blocks, "depth": self.depth,
                "ports": self.num_ports}


class CircuitBuilder:
    """Records a circuit's construction; :meth:`build` flattens it (module docstring)."""

    def __init__(self) -> None:
        self._num_wires = 0
        self._slots: list[tuple[object, tuple[int, ...]]] = []  # (gate name or Circuit, wires)
        self._ports: list[int] = []

    def _new_wires(self, n: int) -> tuple[int, ...]:
        ws = tuple(range(self._num_wires, self._num_wires + n))
        self._num_wires += n
        return ws

    def add_inputs(self, n: int) -> tuple[int, ...]:
        """``n`` new bit wires, each on a new port."""
        if n < 0:
            raise ValueError("Number of bits must be non-negative.")
        ws = self._new_wires(n)
        self._ports.extend(ws)
        return ws

    def add_outputs(self, wires: Iterable[int]) -> None:
        for w in wires:
            if not 0 <= w < self._num_wires:
                raise ValueError(f"Invalid wire index {w}.")
            self._ports.append(int(w))

    def apply(self, block, *wires: int) -> tuple[int, ...]:
        """Apply a gate (a name in :data:`gates.GATES`) or a :class:`Circuit` to ``wires``:
        they go to its first ports, every further port gets a new wire; returns the new
        wires in port order."""
        if isinstance(block, Circuit):
            nports = block.num_ports
        else:
            if block not in gates.GATES:
                raise ValueError(f"unknown gate {block!r}")
            nports = gates.GATES[block].ndim
        if len(wires) > nports:
            raise ValueError(f"Cannot apply block: {len(wires)} wires selected, {nports} ports available.")
        for w in wires:
            if not 0 <= w < self._num_wires:
                raise ValueError(f"Invalid wire {w} selected.")
        new = self._new_wires(nports - len(wires))
        self._slots.append((block, tuple(wires) + new))
        return new

    def build(self, name: str) -> Circuit:
        W = self._num_wires
        # 1. wires joined through a sub-circuit (an inner wire on several of its ports)
        parent = list(range(W))

        def find(x: int) -> int:
            while parent[x] != x:
                parent[x] = parent[parent[x]]
                x = parent[x]
            return x

        for block, ws in self._slots:
            if isinstance(block, Circuit):
                first: dict[int, int] = {}
                for port, inner in enumerate(block.wiring.out_wires):
                    if inner in first:
                        a, b = find(first[inner]), find(ws[port])
                        if a != b:
                            parent[max(a, b)] = min(a, b)
                    else:
                        first[inner] = ws[port]
        remap = [0] * W
        count = 0
        for w in range(W):
            r = find(w)
            if r != w:
                remap[w] = remap[r]
            else:
                remap[w] = count
                count += 1
        # 2. inner wires of sub-circuits not on their ports, in slot order
        inner_maps: list[list[int] | None] = []
        for block, ws in self._slots:
            if not isinstance(block, Circuit):
                inner_maps.append(None)
                continue
            on_port = {inner: ws[port] for port, inner in enumerate(block.wiring.out_wires)}
            m = []
            for inner in range(block.wiring.num_wires):
                if inner in on_port:
                    m.append(remap[on_port[inner]])
                else:
                    m.append(count)
                    count += 1
            inner_maps.append(m)
        # 3. slots: gates in place, sub-circuits expanded
        slot_wires: list[tuple[int, ...]] = []
        blocks: list[np.ndarray] = []
        kinds: list[str] = []
        depth = 0
        for (block, ws), m in zip(self._slots, inner_maps):
            if m is None:
                slot_wires.append(tuple(remap[w] for w in ws))
                blocks.append(gates.GATES[block])
                kinds.append(block)
            else:
                depth = max(depth, block.depth + 1)
                slot_wires.extend(tuple(m[w] for w in sws) for sws in block.wiring.slot_wires_list)
                blocks.extend(block.diagram.blocks)
                kinds.extend(block.diagram.leaf_kinds)
        wiring = Wiring((2,) * count, slot_wires, tuple(remap[w] for w in self._ports))
        return Circuit(name, Diagram(wiring, tuple(blocks), tuple(kinds)), W, len(self._slots), depth)


def _half_adder() -> Circuit:
    """``bincirc.py:91-102``: ports a, b, s = a xor b, c = a and b."""
    c = CircuitBuilder()
    a, b = c.add_inputs(2)
    (s,) = c.apply("xor_", a, b)
    (carry,) = c.apply("and_", a, b)
    c.add_outputs([s, carry])
    return c.build("half_adder")


def _full_adder() -> Circuit:
    """``bincirc.py:105-119``: ports a, b, c_in, s, c_out."""
    c = CircuitBuilder()
    a, b, c_in = c.add_inputs(3)
    (x1,) = c.apply("xor_", a, b)
    (x2,) = c.apply("and_", a, b)
    (x3,) = c.apply("and_", x1, c_in)
    (s,) = c.apply("xor_", x1, c_in)
    (c_out,) = c.apply("or_", x2, x3)
    c.add_outputs([s, c_out])
    return c.build("full_adder")


half_adder: Circuit = _half_adder()
full_adder: Circuit = _full_adder()


def rc_adder(num_bits: int) -> Circuit:
    """
    Ripple-carry adder (``bincirc.py:122-140``): ports ``a[0..n-1]``, ``b[0..n-1]``,
    ``c_in``, then ``s[0..n-1]`` and the final carry; bit i has weight 2^i.  Its relation
    holds exactly the tuples with ``a + b + c_in == s + 2^n * c_out``.
    """
    if not isinstance(num_bits, int) or num_bits <= 0:
        raise ValueError("Number of bits must be positive.")
    c = CircuitBuilder()
    a = c.add_inputs(num_bits)
    b = c.add_inputs(num_bits)
    (carry,) = c.add_inputs(1)
    for i in range(num_bits):
        s, carry = c.apply(full_adder, a[i], b[i], carry)
        c.add_outputs([s])
    c.add_outputs([carry])
    return c.build("rc_adder")


def wallace_columns(num_bits: int) -> tuple[Circuit, list[int]]:
    """The Wallace-tree multiplier and the weight of each of its output ports."""
    if not isinstance(num_bits, int) or num_bits <= 0:
        raise ValueError("Number of bits must be positive.")
    c = CircuitBuilder()
    x = c.add_inputs(num_bits)
    y = c.add_inputs(num_bits)
    # partial products by weight; dict order (first insertion) is the column order the
    # reduction visits and the order of the output ports
    columns: dict[int, list[int]] = {}
    for i in range(num_bits):
        for j in range(num_bits):
            columns.setdefault(i + j, []).extend(c.apply("and_", x[i], y[j]))
    while any(len(col) > 1 for col in columns.values()):
        reduced: dict[int, list[int]] = {}
        for weight, col in columns.items():
            triples, rest = divmod(len(col), 3)
            for t in range(triples):
                s, carry = c.apply(full_adder, *col[3 * t : 3 * t + 3])
                reduced.setdefault(weight, []).append(s)
                reduced.setdefault(weight + 1, []).append(carry)
            if rest == 2:
                s, carry = c.apply(half_adder, *col[-2:])
                reduced.setdefault(weight, []).append(s)
                reduced.setdefault(weight + 1, []).append(carry)
            elif rest == 1:
                reduced.setdefault(weight, []).append(col[-1])
        columns = reduced
    c.add_outputs(col[0] for col in columns.values())
    return c.build("wallace_multiplier"), list(columns.keys())


def wallace_multiplier(num_bits: int) -> Circuit:
    """
    Wallace-tree multiplier (``bincirc.py:143-177``): ports ``x[0..n-1]``, ``y[0..n-1]``,
    then one output per column of the reduced partial-product tree, in the order the
    reduction first created the columns (weights 0, 1, 2, ...; for n >= 3 the tree
    carries into columns above 2n-1, whose bits are always 0).
    """
    return wallace_columns(num_bits)[0]


def constrain(circuit: Circuit, values: dict[int, int]) -> Diagram:
    """
    The circuit with some ports fixed: a ``bit_0`` / ``bit_1`` leaf on the wire of each
    port in ``values`` (port index -> 0 or 1), appended after the circuit's slots; the
    remaining ports, in order, are the out ports.  Contracting it gives the relation of
    the free ports under the constraint (e.g. the factor pairs of a product).
    """
    w = circuit.wiring
    slots = list(w.slot_wires_list)
    blocks = list(circuit.diagram.blocks)
    kinds = list(circuit.diagram.leaf_kinds)
    for port, v in sorted(values.items()):
        if not 0 <= port < len(w.out_wires):
            raise ValueError(f"Invalid port {port}.")
        if v not in (0, 1):
            raise ValueError(f"Port values are bits, got {v!r}.")
        slots.append((w.out_wires[port],))
        blocks.append(gates.bit_1 if v else gates.bit_0)
        kinds.append("bit_1" if v else "bit_0")
    out = tuple(wire for port, wire in enumerate(w.out_wires) if port not in values)
    return Diagram(Wiring(w.wire_types, slots, out), tuple(blocks), tuple(kinds))


def adder_relation(num_bits: int) -> np.ndarray:
    """The dense relation of :func:`rc_adder` from its arithmetic: a (2,)*(3n+2) uint8
    tensor over its ports."""
    n = num_bits
    a = np.arange(1 << n, dtype=np.int64)
    A, B, C = np.meshgrid(a, a, np.arange(2, dtype=np.int64), indexing="ij")
    S = A + B + C
    ports = [(A >> i) & 1 for i in range(n)] + [(B >> i) & 1 for i in range(n)] + [C]
    ports += [(S >> i) & 1 for i in range(n)] + [(S >> n) & 1]
    return _dense(ports)


def multiplier_relation(num_bits: int) -> np.ndarray:
    """The dense relation of :func:`wallace_multiplier` from its arithmetic."""
    n = num_bits
    _, weights = wallace_columns(n)
    a = np.arange(1 << n, dtype=np.int64)
    X, Y = np.meshgrid(a, a, indexing="ij")
    P = X * Y
    ports = [(X >> i) & 1 for i in range(n)] + [(Y >> i) & 1 for i in range(n)]
    ports += [(P >> w) & 1 for w in weights]
    return _dense(ports)


def _dense(port_bits: Sequence[np.ndarray]) -> np.ndarray:
    r = len(port_bits)
    flat = np.zeros(port_bits[0].shape, dtype=np.int64)
    for i, b in enumerate(port_bits):
        flat |= b.astype(np.int64) << (r - 1 - i)
    t = np.zeros(1 << r, dtype=np.uint8)
    t[flat.ravel()] = 1
    return t.reshape((2,) * r)
