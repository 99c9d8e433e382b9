"""
Device-resident Boolean relations: the drop-in counterpart of the reference's ``FinRel``
(``tensorsat/lang/fin_rel.py:162-493``) and of the ``Box`` protocol it implements
(``tensorsat/diagrams.py:693-1031``).

A :class:`DeviceRel` holds its tensor in GPU memory as 0/1 bytes.  Every axis of extent
``d`` occupies ``ceil(log2 d)`` bit axes on the device (entries beyond ``d`` are zero and
stay zero through every operation, since each output entry at a padded index multiplies
an operand entry at that padded index); the logical tensor (``.tensor``) is the numpy
array the reference would hold.  All arithmetic runs in ``libtensorsat_amd.so``:

=================  ==============================  =====================================
reference          this class                      kernel
=================  ==============================  =====================================
``_contract2``     :meth:`DeviceRel._contract2`    ``ts_contract2`` (+``ts_diagonal``,
                                                   ``ts_rewire`` for repeated wires)
``_rewire``        :meth:`DeviceRel._rewire`       ``ts_rewire``
``_spider``        :meth:`DeviceRel._spider`       ``ts_spider``
``_union``         :meth:`DeviceRel._union`        ``ts_union``
``_stack``         :meth:`DeviceRel._stack`        ``ts_stack``
``_getitem``       :meth:`DeviceRel._getitem`      ``ts_select``
``_as_scalar``     :meth:`DeviceRel._as_scalar`    ``ts_as_scalar``
=================  ==============================  =====================================

The epilogue mode is ``"GT0"`` (``count > 0``, the relational meaning) by default;
``with DeviceRel.mode_scope("MOD256"):`` reproduces the reference's numpy arithmetic,
which keeps ``(count mod 256) != 0`` (``fin_rel.py:279,287,392``), including the axis
order of ``rewire_array`` (``_utils/misc.py:104-123``) when ports are discarded or
repeated: its slices are added without a transpose, so the kept ports that occur once
land on the output positions in increasing port order (:meth:`DeviceRel._rewire`).
"""

from __future__ import annotations

import contextlib
import ctypes
import itertools
import math
from collections.abc import Callable, Iterable, Mapping, Sequence
from typing import Any

import numpy as np
import xxhash

from . import _native as N

MAX_DEVICE_BITS = 27


class FinSet:
    """Finite set ``{0, ..., size-1}`` (``fin_rel.py:75-118``); interned by size."""

    _store: dict[int, "FinSet"] = {}
    __slots__ = ("_size",)

    def __new__(cls, size: int) -> "FinSet":
        if not isinstance(size, (int, np.integer)) or isinstance(size, bool):
            raise TypeError("Finite set size must be an int.")
        if size <= 0:
            raise ValueError("Finite set size must be strictly positive.")
        size = int(size)
        self = cls._store.get(size)
        if self is None:
            self = object.__new__(cls)
            self._size = size
            cls._store[size] = self
        return self

    @property
    def size(self) -> int:
        return self._size

    @property
    def tensor_dim(self) -> int:
        return self._size

    def __repr__(self) -> str:
        return f"FinSet({self._size})"


bit = FinSet(2)


def _bits(d: int) -> int:
    return (d - 1).bit_length()


def _sizes(x) -> tuple[int, ...]:
    if isinstance(x, (int, np.integer)):
        return (int(x),)
    if isinstance(x, FinSet):
        return (x.size,)
    return tuple(int(s) if isinstance(s, (int, np.integer)) else s.tensor_dim for s in x)


def _pad(arr: np.ndarray) -> np.ndarray:
    """Logical array -> flat padded bytes (bit axes, MSB first; extent-1 axes vanish)."""
    arr = np.asarray(arr, dtype=np.uint8)
    padded_shape = tuple(1 << _bits(d) for d in arr.shape)
    if padded_shape == arr.shape:
        return np.ascontiguousarray(arr).reshape(-1)
    out = np.zeros(padded_shape, dtype=np.uint8)
    out[tuple(slice(0, d) for d in arr.shape)] = arr
    return out.reshape(-1)


def _unpad(flat: np.ndarray, shape: tuple[int, ...]) -> np.ndarray:
    padded_shape = tuple(1 << _bits(d) for d in shape)
    a = flat.reshape(padded_shape) if padded_shape else flat.reshape(())
    if padded_shape == shape:
        return np.array(a, order="C")  # (np.ascontiguousarray would turn 0-d into 1-d)
    return np.array(a[tuple(slice(0, d) for d in shape)], order="C")


def _torch():
    import torch

    return torch


def _empty(nbytes: int):
    torch = _torch()
    # at least 16 bytes so vector paths never read past the allocation
    return torch.empty(max(16, nbytes), dtype=torch.uint8, device="cuda")


class DeviceRel:
    """A finite relation held on the GPU as a Boolean tensor (``FinRel`` counterpart)."""

    _mode: str = "GT0"
    __slots__ = ("_data", "_shape", "_name", "_hash")

    # ---- construction --------------------------------------------------------------

    @classmethod
    def _new(cls, data, shape: tuple[int, ...], name: str | None = None) -> "DeviceRel":
        self = object.__new__(cls)
        self._data = data
        self._shape = tuple(int(d) for d in shape)
        self._name = name
        return self

    def __new__(cls, tensor: np.ndarray, *, name: str | None = None) -> "DeviceRel":
        """Relation from a uint8 0/1 tensor (``fin_rel.py:328-340``)."""
        tensor = np.asarray(tensor)
        if tensor.dtype != np.uint8:
            raise TypeError("Boolean tensors must be numpy uint8 arrays.")
        if not np.all(tensor <= 1):
            raise ValueError("Values in a Boolean tensor must be 0 or 1.")
        if any(d == 0 for d in tensor.shape):
            raise ValueError("Zero dimension in shape.")
        return cls.from_numpy(tensor, name=name)

    @classmethod
    def from_numpy(cls, tensor: np.ndarray, *, name: str | None = None) -> "DeviceRel":
        torch = _torch()
        N.lib()
        flat = _pad(tensor)
        if not flat.flags.writeable:
            flat = flat.copy()
        if int(np.log2(max(1, flat.size))) > MAX_DEVICE_BITS:
            raise NotImplementedError("tensor exceeds 2^27 device entries")
        data = _empty(flat.size)
        data[: flat.size].copy_(torch.from_numpy(flat))
        return cls._new(data, tensor.shape, name)

    @classmethod
    def from_set(cls, shape, points: Iterable, *, name: str | None = None) -> "DeviceRel":
        """``fin_rel.py:168-190``."""
        shape = _sizes(shape)
        if any(d <= 0 for d in shape):
            raise ValueError("Zero dimension in shape.")
        data = np.zeros(shape, dtype=np.uint8)
        for point in points:
            if isinstance(point, (int, np.integer)):
                point = (int(point),)
            point = tuple(point)
            if len(point) != len(shape):
                raise ValueError(f"Length of {point = } is invalid for {shape = }.")
            if not all(0 <= i < d for i, d in zip(point, shape)):
                raise ValueError(f"Values of {point = } are invalid for {shape = }.")
            data[point] = 1
        return cls.from_numpy(data, name=name)

    @classmethod
    def from_mapping(cls, input_shape, output_shape, mapping: Mapping, *, name: str | None = None) -> "DeviceRel":
        """``fin_rel.py:192-213``."""
        input_shape = _sizes(input_shape)
        output_shape = _sizes(output_shape)
        pts = []
        for k, v in mapping.items():
            v = (v,) if isinstance(v, (int, np.integer)) else tuple(v)
            pts.append(tuple(k) + v)
        rel = cls.from_set(input_shape + output_shape, pts, name=name)
        if len(mapping) != math.prod(input_shape):
            raise ValueError("Mapping does not cover the entire input space.")
        return rel

    @classmethod
    def singleton(cls, shape, point, *, name: str | None = None) -> "DeviceRel":
        return cls.from_mapping((), shape, {(): point}, name=name)

    @classmethod
    def from_callable(cls, input_shape, output_shape, func: Callable, *, name: str | None = None) -> "DeviceRel":
        input_shape = _sizes(input_shape)
        output_shape = _sizes(output_shape)
        mapping = {idx: func(*idx) for idx in np.ndindex(input_shape)}
        return cls.from_mapping(input_shape, output_shape, mapping, name=name)

    # ---- properties -----------------------------------------------------------------

    @property
    def name(self) -> str | None:
        return self._name

    @property
    def shape(self) -> tuple[FinSet, ...]:
        return tuple(FinSet(d) for d in self._shape)

    @property
    def tensor_shape(self) -> tuple[int, ...]:
        return self._shape

    @property
    def num_ports(self) -> int:
        return len(self._shape)

    @property
    def ports(self) -> range:
        return range(len(self._shape))

    @property
    def device_bits(self) -> int:
        return sum(_bits(d) for d in self._shape)

    @property
    def data(self):
        """The device bytes (torch uint8), padded layout."""
        return self._data

    @property
    def tensor(self) -> np.ndarray:
        """Host copy of the logical tensor (read-only)."""
        n = 1 << self.device_bits
        flat = self._data[:n].cpu().numpy()
        t = _unpad(flat, self._shape)
        t.setflags(write=False)
        return t

    # ---- mode ---------------------------------------------------------------------

    @classmethod
    @contextlib.contextmanager
    def mode_scope(cls, mode: str):
        if mode not in N.MODES:
            raise ValueError(f"unknown mode {mode!r}")
        old = DeviceRel._mode
        DeviceRel._mode = mode
        try:
            yield
        finally:
            DeviceRel._mode = old

    # ---- bit-level helpers ----------------------------------------------------------

    def _axis_bits(self) -> list[int]:
        return [_bits(d) for d in self._shape]

    @staticmethod
    def _bitwires(wires: Sequence[int], extents: Sequence[int], ids: dict) -> list[int]:
        out = []
        for w, d in zip(wires, extents):
            for k in range(_bits(d)):
                key = (w, k)
                if key not in ids:
                    ids[key] = len(ids)
                out.append(ids[key])
        return out

    def _diag(self, wires: Sequence[int]) -> tuple["DeviceRel", list[int]]:
        """Take the diagonal over repeated wires; returns (box, unique wires)."""
        uniq: list[int] = []
        for w in wires:
            if w not in uniq:
                uniq.append(w)
        if len(uniq) == len(wires):
            return self, list(wires)
        ext = {}
        for w, d in zip(wires, self._shape):
            if ext.setdefault(w, d) != d:
                raise ValueError("A wire joins ports of different sizes.")
        out_shape = tuple(ext[w] for w in uniq)
        # bit positions: out bit axis index for each input bit axis
        out_bit_index = {}
        pos = 0
        for w in uniq:
            for k in range(_bits(ext[w])):
                out_bit_index[(w, k)] = pos
                pos += 1
        axis_map = [out_bit_index[(w, k)] for w, d in zip(wires, self._shape) for k in range(_bits(d))]
        res = _empty(1 << pos)
        N.check(N.lib().ts_diagonal(N.dptr(self._data), self.device_bits, N.i32arr(axis_map), N.dptr(res), pos,
                                    N.stream_ptr()), "ts_diagonal")
        return DeviceRel._new(res, out_shape), uniq

    # ---- Box protocol ----------------------------------------------------------------

    @classmethod
    def class_join(cls, bases) -> type:
        return cls

    @classmethod
    def can_be_contracted(cls) -> bool:
        return True

    @classmethod
    def contract2(cls, lhs: "DeviceRel", lhs_wires, rhs: "DeviceRel", rhs_wires, out_wires=None) -> "DeviceRel":
        """``Box.contract2`` (``diagrams.py:811-844``): validation, default out wires."""
        lhs_wires, rhs_wires = list(lhs_wires), list(rhs_wires)
        if len(lhs_wires) != len(lhs.shape):
            raise ValueError(
                f"Number of wires in lhs ({len(lhs_wires)}) does not match"
                f" the number of ports in lhs shape ({len(lhs.shape)})."
            )
        if len(rhs_wires) != len(rhs.shape):
            raise ValueError(
                f"Number of wires in rhs ({len(rhs_wires)}) does not match"
                f" the number of ports in rhs shape ({len(rhs.shape)})."
            )
        if out_wires is None:
            out_wires = sorted(set(lhs_wires).symmetric_difference(rhs_wires))
        else:
            out_wires = list(out_wires)
            if set(out_wires) - set(lhs_wires) - set(rhs_wires):
                raise ValueError("Every output wire must appear in LHR or RHS.")
        return cls._contract2(lhs, lhs_wires, rhs, rhs_wires, out_wires)

    @classmethod
    def _contract2(cls, lhs: "DeviceRel", lhs_wires, rhs: "DeviceRel", rhs_wires, out_wires) -> "DeviceRel":
        """``FinRel._contract2`` (``fin_rel.py:255-288``) on the device."""
        lhs, lw = lhs._diag(list(lhs_wires))
        rhs, rw = rhs._diag(list(rhs_wires))
        ext: dict[int, int] = {}
        for w, d in itertools.chain(zip(lw, lhs._shape), zip(rw, rhs._shape)):
            if ext.setdefault(w, d) != d:
                raise ValueError("A wire joins ports of different sizes.")
        out_wires = list(out_wires)
        uniq_out: list[int] = []
        for w in out_wires:
            if w not in uniq_out:
                uniq_out.append(w)
        if len(uniq_out) != len(out_wires) and DeviceRel._mode == "MOD256":
            # fin_rel.py:281-287: the einsum keeps sorted(set(out_wires)) and rewire_array
            # places the repeated wires relative to that order.
            uniq_out = sorted(uniq_out)
        ids: dict = {}
        bl = cls._bitwires(lw, lhs._shape, ids)
        br = cls._bitwires(rw, rhs._shape, ids)
        out_shape = tuple(ext[w] for w in uniq_out)
        bo = cls._bitwires(uniq_out, out_shape, ids)
        if max(len(bl), len(br), len(bo)) > MAX_DEVICE_BITS:
            raise NotImplementedError("operand exceeds 2^27 device entries")
        res = _empty(1 << len(bo))
        N.check(N.lib().ts_contract2(
            N.dptr(lhs._data), len(bl), N.i32arr(bl), N.dptr(rhs._data), len(br), N.i32arr(br),
            N.dptr(res), len(bo), N.i32arr(bo), N.MODES[DeviceRel._mode], N.stream_ptr()), "ts_contract2")
        box = DeviceRel._new(res, out_shape)
        if len(uniq_out) != len(out_wires):
            box = box._rewire([uniq_out.index(w) for w in out_wires])
        return box

    @classmethod
    def spider(cls, t, num_ports: int) -> "DeviceRel":
        """``Box.spider`` (``diagrams.py:777-795``)."""
        if not isinstance(num_ports, (int, np.integer)):
            raise TypeError("Number of ports must be an int.")
        if num_ports <= 0:
            raise ValueError("Number of ports must be strictly positive.")
        return cls._spider(t, int(num_ports))

    @classmethod
    def _spider(cls, t, num_ports: int) -> "DeviceRel":
        """``fin_rel.py:290-298``."""
        d = t.tensor_dim if isinstance(t, FinSet) else int(t)
        b = _bits(d)
        if num_ports * b > MAX_DEVICE_BITS:
            raise NotImplementedError("spider exceeds 2^27 device entries")
        res = _empty(1 << (num_ports * b))
        N.check(N.lib().ts_spider(N.dptr(res), num_ports, b, d, N.stream_ptr()), "ts_spider")
        return DeviceRel._new(res, (d,) * num_ports)

    @classmethod
    def scalar(cls, value: bool) -> "DeviceRel":
        return cls._scalar(value)

    @classmethod
    def _scalar(cls, value: bool) -> "DeviceRel":
        return cls.from_numpy(np.array(int(bool(value)), dtype=np.uint8))

    @classmethod
    def prod(cls, boxes: Iterable["DeviceRel"]) -> "DeviceRel":
        """``Box.prod`` (``diagrams.py:719-746``): left fold of outer products."""
        boxes = tuple(boxes)
        if not boxes:
            raise ValueError("Products must involve at least one box.")
        res = boxes[0]
        for box in boxes[1:]:
            nl, nr = len(res.shape), len(box.shape)
            res = cls._contract2(res, range(nl), box, range(nl, nl + nr), range(nl + nr))
        return res

    @classmethod
    def union(cls, boxes: Iterable["DeviceRel"]) -> "DeviceRel":
        boxes = tuple(boxes)
        if not boxes:
            raise ValueError("Cannot take an empty union: no shape info available.")
        res = boxes[0]
        for box in boxes[1:]:
            if box._shape != res._shape:
                raise ValueError("Shapes of boxes in a union must coincide.")
            res = res._union(box)
        return res

    def _union(self, other: "DeviceRel") -> "DeviceRel":
        n = 1 << self.device_bits
        res = _empty(n)
        N.check(N.lib().ts_union(N.dptr(self._data), N.dptr(other._data), N.dptr(res), n, N.stream_ptr()),
                "ts_union")
        return DeviceRel._new(res, self._shape)

    @classmethod
    def stack(cls, boxes: Iterable["DeviceRel"], port: int = 0) -> "DeviceRel":
        """``Box.stack`` (``diagrams.py:883-901``)."""
        boxes = tuple(boxes)
        if not boxes:
            raise ValueError("Need at least one box to stack.")
        shape = boxes[0]._shape
        if port not in range(len(shape) + 1):
            raise ValueError(f"Invalid port index {port} for stacking.")
        for b in boxes[1:]:
            if b._shape != shape:
                raise ValueError("Boxes to stack must have the same shape.")
        return cls._stack(boxes, port)

    @classmethod
    def _stack(cls, boxes: tuple["DeviceRel", ...], port: int = 0) -> "DeviceRel":
        shape = boxes[0]._shape
        nb = len(boxes)
        axis_bits = _bits(nb)
        in_bits = boxes[0].device_bits
        bit_port = sum(_bits(d) for d in shape[:port])
        if in_bits + axis_bits > MAX_DEVICE_BITS:
            raise NotImplementedError("stack exceeds 2^27 device entries")
        res = _empty(1 << (in_bits + axis_bits))
        ptrs = (ctypes.c_void_p * nb)(*[b._data.data_ptr() for b in boxes])
        N.check(N.lib().ts_stack(ptrs, nb, in_bits, bit_port, axis_bits, N.dptr(res), N.stream_ptr()), "ts_stack")
        return DeviceRel._new(res, shape[:port] + (nb,) + shape[port:])

    def rewire(self, out_ports: Sequence[int]) -> "DeviceRel":
        """``Box.rewire`` (``diagrams.py:922-939``)."""
        out_ports = list(out_ports)
        if not all(isinstance(p, (int, np.integer)) and 0 <= p < self.num_ports for p in out_ports):
            raise ValueError("Output ports must be valid for this box.")
        return self._rewire(out_ports)

    def _rewire(self, out_ports: Sequence[int]) -> "DeviceRel":
        """``fin_rel.py:379-393``: permute / duplicate / discard (existential sum).

        GT0 computes the intended map.  MOD256 follows the reference literally: outside
        a pure permutation, ``rewire_array`` (``_utils/misc.py:104-123``) adds slices of
        the input whose free axes are in increasing port order into output slices whose
        free axes are in out-port order, so the ports that occur once in ``out_ports``
        are taken in increasing order (:meth:`_literal_ports`)."""
        out_ports = [int(p) for p in out_ports]
        out_shape = tuple(self._shape[p] for p in out_ports)
        ports = out_ports
        if DeviceRel._mode == "MOD256" and not (len(set(out_ports)) == len(out_ports) == len(self._shape)):
            ports = self._literal_ports(out_ports)
        start = []
        acc = 0
        for d in self._shape:
            start.append(acc)
            acc += _bits(d)
        bit_ports = [start[p] + k for p in ports for k in range(_bits(self._shape[p]))]
        if len(bit_ports) > MAX_DEVICE_BITS:
            raise NotImplementedError("rewire exceeds 2^27 device entries")
        res = _empty(1 << len(bit_ports))
        N.check(N.lib().ts_rewire(N.dptr(self._data), self.device_bits, N.dptr(res), len(bit_ports),
                                  N.i32arr(bit_ports), N.MODES[DeviceRel._mode], N.stream_ptr()), "ts_rewire")
        return DeviceRel._new(res, out_shape)

    def _literal_ports(self, out_ports: list[int]) -> list[int]:
        """The ports ``rewire_array`` actually reads for each output position: the
        ports occurring once in ``out_ports`` are reassigned, position by position, in
        increasing order; repeated ports stay (diagonals).  numpy adds a slice of
        extents ``shape[sorted port]`` into one of extents ``shape[out port]``, so
        unequal extents fail as they do there (extent 1 would broadcast: not
        supported)."""
        counts: dict[int, int] = {}
        for p in out_ports:
            counts[p] = counts.get(p, 0) + 1
        single = [q for q, p in enumerate(out_ports) if counts[p] == 1]
        ports = list(out_ports)
        for q, s in zip(single, sorted(out_ports[q] for q in single)):
            want, have = self._shape[out_ports[q]], self._shape[s]
            if have != want:
                if have == 1:
                    raise NotImplementedError("rewire_array would broadcast an extent-1 port")
                raise ValueError(f"operands could not be broadcast together: port {s} of extent {have} "
                                 f"lands on a port of extent {want}")
            ports[q] = s
        return ports

    def __getitem__(self, idxs) -> "DeviceRel":
        """``Box.__getitem__`` (``diagrams.py:972-996``)."""
        if not isinstance(idxs, tuple):
            raise TypeError("Port indexing needs a tuple.")
        for idx in idxs:
            if not isinstance(idx, (int, np.integer)):
                if not isinstance(idx, slice):
                    raise TypeError("Indices must be ints or full slices.")
                if idx != slice(None, None, None):
                    raise NotImplementedError("Non-trivial slicing is not implemented.")
        return self._getitem(idxs)

    def _getitem(self, idxs) -> "DeviceRel":
        idxs = tuple(idxs) + (slice(None),) * (self.num_ports - len(idxs))
        fixed = []
        out_shape = []
        for idx, d in zip(idxs, self._shape):
            b = _bits(d)
            if isinstance(idx, slice):
                fixed.extend([-1] * b)
                out_shape.append(d)
            else:
                v = int(idx)
                if not 0 <= v < d:
                    raise IndexError("port index out of range")
                fixed.extend((v >> (b - 1 - k)) & 1 for k in range(b))
        res = _empty(1 << sum(1 for f in fixed if f < 0))
        N.check(N.lib().ts_select(N.dptr(self._data), self.device_bits, N.i32arr(fixed), N.dptr(res),
                                  N.stream_ptr()), "ts_select")
        return DeviceRel._new(res, tuple(out_shape))

    def __add__(self, other: "DeviceRel") -> "DeviceRel":
        """``Box.__add__`` (``diagrams.py:965-970``)."""
        if self._shape != other._shape:
            raise ValueError("Boxes in a union must have the same shape")
        return self._union(other)

    def __bool__(self) -> bool:
        """``Box.__bool__`` (``diagrams.py:949-958``)."""
        if self._shape:
            raise ValueError("Box is not a scalar.")
        return self._as_scalar()

    def _as_scalar(self) -> bool:
        """``fin_rel.py:472-473``."""
        out = N.i32(0)
        N.check(N.lib().ts_as_scalar(N.dptr(self._data), 1 << self.device_bits, ctypes.byref(out), N.stream_ptr()),
                "ts_as_scalar")
        return bool(out.value)

    # ---- host-side views (as FinRel) -------------------------------------------------

    def to_set(self) -> frozenset:
        t = self.tensor
        return frozenset(tuple(int(i) for i in idx) for idx in np.argwhere(t))

    def _function_matrix(self, input_ports):
        ports = self.ports
        if not all(p in ports for p in input_ports):
            raise ValueError("Invalid input ports.")
        if len(input_ports) != len(set(input_ports)):
            raise ValueError("Input ports cannot be repeated.")
        output_ports = tuple(p for p in ports if p not in input_ports)
        t = np.transpose(self.tensor, tuple(input_ports) + output_ports)
        ni = math.prod(self._shape[i] for i in input_ports)
        no = math.prod(self._shape[o] for o in output_ports)
        return output_ports, t.reshape(ni, no)

    def is_function_graph(self, input_ports) -> bool:
        _, m = self._function_matrix(tuple(input_ports))
        return bool(np.all(np.count_nonzero(m, axis=1) == 1))

    def to_mapping(self, input_ports) -> dict:
        input_ports = tuple(input_ports)
        output_ports, m = self._function_matrix(input_ports)
        if not np.all(np.count_nonzero(m, axis=1) == 1):
            raise ValueError("Relation is not a function graph with the given input ports.")
        in_sizes = tuple(self._shape[i] for i in input_ports)
        out_sizes = tuple(self._shape[o] for o in output_ports)
        out_points = list(np.ndindex(out_sizes))
        return {pt: out_points[int(np.argmax(row))] for pt, row in zip(np.ndindex(in_sizes), m)}

    def __eq__(self, other: Any) -> bool:
        if not isinstance(other, DeviceRel):
            return NotImplemented
        if self is other:
            return True
        return self._shape == other._shape and np.array_equal(self.tensor, other.tensor)

    def __hash__(self) -> int:
        """xxh64 of the logical tensor bytes, as ``fin_rel.py:487-493``."""
        try:
            return self._hash
        except AttributeError:
            self._hash = h = xxhash.xxh64(self.tensor.tobytes()).intdigest()
            return h

    def __repr__(self) -> str:
        if self._name is None:
            return f"<DeviceRel: {len(self._shape)} ports>"
        return f"<DeviceRel: {len(self._shape)} ports, named {self._name!r}>"
