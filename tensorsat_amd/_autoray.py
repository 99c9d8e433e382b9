"""
Array-backend functions for contraction executors, the counterpart of the reference's
``tensorsat/_autoray.py:31-80``: the pairwise ``einsum`` a tree executor calls per node,
``transpose``, ``shape`` and ``stack``, all dispatching to the box class of the operands
(:class:`tensorsat_amd.rel.DeviceRel` on the GPU).

Wire numbering follows the reference (``_autoray.py:54-60``): the wire of a letter is the
position of its first occurrence in ``lhs + rhs``.  Single-operand einsum raises
``NotImplementedError`` as in the reference (``_autoray.py:42-50``).
"""

from __future__ import annotations

from collections.abc import Sequence


def _class_join(boxes):
    cls = type(boxes[0])
    for b in boxes[1:]:
        if type(b) is not cls:
            raise TypeError("Boxes of different classes cannot be combined.")
    return cls


def einsum(contraction: str, /, lhs, rhs=None):
    """Contract boxes using einsum notation (``_autoray.py:31-62``)."""
    if not isinstance(contraction, str):
        raise TypeError("einsum expects a contraction string")
    _input_wires, _out_wires = contraction.split("->")
    if rhs is None:
        raise NotImplementedError("This case needs to be handled specially.")
    box_class = _class_join([lhs, rhs])
    _lhs_wires, _rhs_wires = _input_wires.split(",")
    char_idxs = {letter: idx for idx, letter in reversed(list(enumerate(_lhs_wires + _rhs_wires)))}
    lhs_wires = [char_idxs[c] for c in _lhs_wires]
    rhs_wires = [char_idxs[c] for c in _rhs_wires]
    out_wires = [char_idxs[c] for c in _out_wires]
    return box_class.contract2(lhs, lhs_wires, rhs, rhs_wires, out_wires)


def transpose(box, perm: Sequence[int], /):
    """Rearrange the ports of a box (``_autoray.py:64-67``)."""
    return box.rewire(perm)


def shape(box) -> tuple[int, ...]:
    """Tensor shape of a box (``_autoray.py:69-73``)."""
    ts = getattr(box, "tensor_shape", None)
    if ts is None:
        raise TypeError("Autoray shape is only supported by tensor-like boxes.")
    return tuple(ts)


def stack(boxes: Sequence, axis: int = 0):
    """Stack boxes along a new axis (``_autoray.py:75-80``)."""
    boxes = list(boxes)
    box_class = _class_join(boxes)
    return box_class.stack(boxes, axis)
