"""
ORACLE — TEST INFRASTRUCTURE ONLY.  Restatement of ``FinRel`` arithmetic.

Reference anchors (all in ``tensorsat/``):

* ``lang/fin_rel.py:145-159``  ``_to_safe_contraction_dtype``  -> :func:`safe_contraction_dtype`
* ``lang/fin_rel.py:255-288``  ``FinRel._contract2``           -> :func:`contract2`
* ``lang/fin_rel.py:290-298``  ``FinRel._spider``              -> :func:`spider`
* ``lang/fin_rel.py:304-308``  ``FinRel._stack``               -> :func:`stack`
* ``lang/fin_rel.py:342-346``  ``_getitem`` / ``_union``       -> :func:`getitem`, :func:`union`
* ``lang/fin_rel.py:379-393``  ``FinRel._rewire``              -> :func:`rewire`
* ``lang/fin_rel.py:472-473``  ``FinRel._as_scalar``           -> :func:`as_scalar`
* ``_utils/misc.py:104-123``   ``rewire_array``                -> :func:`rewire_array`
* ``diagrams.py:719-746``      ``Box.prod``                    -> :func:`prod`

Epilogue modes:

* ``"MOD256"`` restates the reference's numpy steps literally: widen by the
  contracted size, integer einsum in that dtype (wraps modulo 2^bits), then
  ``np.sign(res, dtype=np.uint8)``, which casts to uint8 *before* the sign.  The value
  is therefore ``(count mod 256) != 0``.
* ``"GT0"`` is ``count > 0`` with an exact int64 count.
* :func:`count_contract2` is the counting semiring over Python ints (object arrays),
  never rounded.

Pinned by ``tests/test_oracle.py``: notebook goldens of ``01-introduction.ipynb``
(cells 6, 11, 42-55, 92, 102-104), the all-ones 256-term divergence closed form,
brute-force ``np.einsum`` in int64 on random small shapes, and the equality of the
literal MOD256 restatement with the closed form ``(count mod 256) != 0``.
"""

from __future__ import annotations

from collections import Counter
from itertools import product
from math import prod as _prod

import numpy as np

MODES = ("GT0", "MOD256")


def _labels(*wire_lists):
    """Map arbitrary wire ids to einsum sublist labels 0..51."""
    ids: dict[int, int] = {}
    for ws in wire_lists:
        for w in ws:
            if w not in ids:
                ids[w] = len(ids)
    if len(ids) > 52:
        raise ValueError("too many distinct wires for numpy einsum")
    return [[ids[w] for w in ws] for ws in wire_lists]


def safe_contraction_dtype(contraction_size: int, *tensors: np.ndarray):
    """``fin_rel.py:145-159``: keep uint8 below 256, else widen to 16/32/64 bits."""
    assert contraction_size >= 0
    if contraction_size < 256:
        return tensors
    if contraction_size < 65536:
        dt = np.dtype("uint16")
    elif contraction_size < 2**32:
        dt = np.dtype("uint32")
    else:
        dt = np.dtype("uint64")
    return tuple(t.astype(dt) for t in tensors)


def rewire_array(a: np.ndarray, out_ports) -> np.ndarray:
    """
    ``_utils/misc.py:104-123``: output ``b`` of shape ``a.shape[p] for p in out_ports``;
    for every value combination of the repeated and discarded ports, the slice of
    ``a`` with those ports fixed is added into the slice of ``b`` with them fixed.

    The two slices keep their remaining axes in their own orders (``a``'s port order,
    ``b``'s out-port order) and are added without transposing, as in the reference.
    When the non-repeated out ports are not in increasing order this differs from a
    true permutation; :func:`rewire_exact` is the intended map.
    """
    out_ports = list(out_ports)
    a_ports = range(a.ndim)
    b = np.zeros(tuple(a.shape[p] for p in out_ports), dtype=a.dtype)
    repeat = sorted(p for p, c in Counter(out_ports).items() if c > 1)
    repeat += sorted(set(a_ports) - set(out_ports))
    pos = {p: i for i, p in enumerate(repeat)}
    for vals in product(*(range(a.shape[p]) for p in repeat)):
        a_idx = tuple(vals[pos[p]] if p in pos else slice(None) for p in a_ports)
        b_idx = tuple(vals[pos[p]] if p in pos else slice(None) for p in out_ports)
        b[b_idx] += a[a_idx]
    return b


def rewire_exact(a: np.ndarray, out_ports, dtype=np.int64) -> np.ndarray:
    """The intended rewire: ``b[i_0..] = sum over discarded ports of a`` with repeated
    out ports on the diagonal (exact integer counts)."""
    out_ports = list(out_ports)
    lab = list(range(a.ndim))
    kept = sorted(set(out_ports))
    s = np.einsum(a.astype(dtype), lab, kept)
    b = np.zeros(tuple(a.shape[p] for p in out_ports), dtype=dtype)
    for idx in np.ndindex(*[a.shape[p] for p in kept]):
        val = s[idx]
        coord = dict(zip(kept, idx))
        b[tuple(coord[p] for p in out_ports)] = val
    return b


def _count_einsum(lhs, lw, rhs, rw, ow, dtype=np.int64):
    ll, rl, ol = _labels(lw, rw, ow)
    return np.einsum(lhs.astype(dtype), ll, rhs.astype(dtype), rl, ol)


def contract2(lhs, lhs_wires, rhs, rhs_wires, out_wires=None, mode: str = "GT0") -> np.ndarray:
    """
    ``FinRel._contract2`` (``fin_rel.py:255-288``) with the default out wires of
    ``Box.contract2`` (``diagrams.py:836-837``: sorted symmetric difference).

    ``mode="MOD256"`` is the literal restatement (widening + wrapped integer einsum +
    ``np.sign(dtype=uint8)``); ``mode="GT0"`` returns ``count > 0``.
    """
    lhs_wires, rhs_wires = list(lhs_wires), list(rhs_wires)
    if len(lhs_wires) != lhs.ndim or len(rhs_wires) != rhs.ndim:
        raise ValueError("Number of wires does not match the number of ports.")
    if out_wires is None:
        out_wires = sorted(set(lhs_wires).symmetric_difference(rhs_wires))
    out_wires = list(out_wires)
    if set(out_wires) - set(lhs_wires) - set(rhs_wires):
        raise ValueError("Every output wire must appear in LHR or RHS.")
    if mode == "GT0":
        if len(set(out_wires)) == len(out_wires):
            cnt = _count_einsum(lhs, lhs_wires, rhs, rhs_wires, out_wires)
            return (cnt > 0).astype(np.uint8)
        uniq = sorted(set(out_wires))
        cnt = _count_einsum(lhs, lhs_wires, rhs, rhs_wires, uniq)
        res = (cnt > 0).astype(np.uint8)
        return rewire_exact(res, [uniq.index(w) for w in out_wires]).astype(np.uint8)
    if mode != "MOD256":
        raise ValueError(f"unknown mode {mode!r}")
    common = set(lhs_wires) & set(rhs_wires)
    csize = _prod(d for d, w in zip(lhs.shape, lhs_wires) if w in common - set(out_wires))
    lt, rt = safe_contraction_dtype(csize, lhs, rhs)
    if len(set(out_wires)) == len(out_wires):
        ll, rl, ol = _labels(lhs_wires, rhs_wires, out_wires)
        res = np.einsum(lt, ll, rt, rl, ol)
        return np.sign(res, dtype=np.uint8)
    einsum_out = sorted(set(out_wires))
    ports = [einsum_out.index(w) for w in out_wires]
    ll, rl, ol = _labels(lhs_wires, rhs_wires, einsum_out)
    res = np.einsum(lt, ll, rt, rl, ol)
    res = rewire_array(res, ports)
    return np.sign(res, dtype=np.uint8)


def contract2_closed_form(lhs, lhs_wires, rhs, rhs_wires, out_wires, mode: str) -> np.ndarray:
    """Closed form of both modes on unique out wires: exact count, then ``>0`` or
    ``mod 256 != 0``.  Used to pin :func:`contract2` and as the kernels' spec."""
    cnt = _count_einsum(lhs, lhs_wires, rhs, rhs_wires, out_wires)
    if mode == "GT0":
        return (cnt > 0).astype(np.uint8)
    return ((cnt % 256) != 0).astype(np.uint8)


def count_contract2(lhs, lhs_wires, rhs, rhs_wires, out_wires) -> np.ndarray:
    """Counting semiring: exact sums of products over Python ints (object arrays)."""
    ll, rl, ol = _labels(lhs_wires, rhs_wires, out_wires)
    a = np.asarray(lhs).astype(object)
    b = np.asarray(rhs).astype(object)
    return np.einsum(a, ll, b, rl, ol)


def rewire(t: np.ndarray, out_ports, mode: str = "MOD256") -> np.ndarray:
    """
    ``FinRel._rewire`` (``fin_rel.py:379-393``): a pure permutation is a transpose;
    otherwise widen by the discarded size, :func:`rewire_array`, ``np.sign(uint8)``.
    ``mode="GT0"`` uses :func:`rewire_exact` and ``count > 0``.
    """
    out_ports = list(out_ports)
    if not all(0 <= p < t.ndim for p in out_ports):
        raise ValueError("Output ports must be valid for this box.")
    if len(set(out_ports)) == len(out_ports) == t.ndim:
        return np.array(np.transpose(t, out_ports), order="C")  # keeps 0-d shapes
    if mode == "GT0":
        return (rewire_exact(t, out_ports) > 0).astype(np.uint8)
    discarded = set(range(t.ndim)) - set(out_ports)
    a = t
    if discarded:
        csize = _prod(t.shape[p] for p in discarded)
        (a,) = safe_contraction_dtype(csize, t)
    res = rewire_array(a, out_ports)
    return np.sign(res, dtype=np.uint8)


def spider(dim: int, num_ports: int) -> np.ndarray:
    """``fin_rel.py:290-298``: ones on the diagonal of a ``(dim,)*num_ports`` tensor."""
    if num_ports <= 0:
        raise ValueError("Number of ports must be strictly positive.")
    t = np.zeros((dim,) * num_ports, dtype=np.uint8)
    for i in range(dim):
        t[(i,) * num_ports] = 1
    return t


def union(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """``fin_rel.py:345-346``."""
    if a.shape != b.shape:
        raise ValueError("Boxes in a union must have the same shape")
    return a | b


def stack(tensors, axis: int = 0) -> np.ndarray:
    """``fin_rel.py:304-308``."""
    return np.stack(list(tensors), axis=axis)


def getitem(t: np.ndarray, idxs) -> np.ndarray:
    """``fin_rel.py:342-343`` with the checks of ``diagrams.py:972-996``."""
    for idx in idxs:
        if isinstance(idx, slice) and idx != slice(None):
            raise NotImplementedError("Non-trivial slicing is not implemented.")
    return np.array(t[tuple(idxs)], order="C")


def as_scalar(t: np.ndarray) -> bool:
    """``fin_rel.py:472-473``."""
    return bool(np.all(t == 1))


def prod(tensors, mode: str = "GT0") -> np.ndarray:
    """``Box.prod`` (``diagrams.py:719-746``): left fold of outer products."""
    tensors = list(tensors)
    if not tensors:
        raise ValueError("Products must involve at least one box.")
    res = tensors[0]
    for t in tensors[1:]:
        nl, nr = res.ndim, t.ndim
        res = contract2(res, range(nl), t, range(nl, nl + nr), range(nl + nr), mode=mode)
    return res
