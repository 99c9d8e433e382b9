"""
A host box class over the oracle (tests only): the Box protocol of the reference
(``diagrams.py:693-1031``) implemented with ``oracle.finrel``, so the contraction classes'
host logic can be checked on the CPU against the same goldens as the device box.
"""

from __future__ import annotations

import numpy as np
import xxhash

from oracle import finrel as F


class OracleRel:
    mode = "GT0"

    def __init__(self, tensor):
        self.tensor = np.array(tensor, dtype=np.uint8, order="C")

    @classmethod
    def from_numpy(cls, t):
        return cls(t)

    @classmethod
    def can_be_contracted(cls):
        return True

    @property
    def tensor_shape(self):
        return self.tensor.shape

    @property
    def shape(self):
        return self.tensor.shape

    @classmethod
    def contract2(cls, lhs, lw, rhs, rw, out=None):
        return cls(F.contract2(lhs.tensor, list(lw), rhs.tensor, list(rw), None if out is None else list(out), cls.mode))

    @classmethod
    def _contract2(cls, lhs, lw, rhs, rw, out):
        return cls(F.contract2(lhs.tensor, list(lw), rhs.tensor, list(rw), list(out), cls.mode))

    def _rewire(self, ports):
        return OracleRel(F.rewire(self.tensor, list(ports), "GT0" if OracleRel.mode == "GT0" else "MOD256"))

    rewire = _rewire

    @classmethod
    def spider(cls, t, n):
        return cls(F.spider(int(t), n))

    @classmethod
    def scalar(cls, v):
        return cls(np.array(int(bool(v)), np.uint8))

    @classmethod
    def prod(cls, boxes):
        return cls(F.prod([b.tensor for b in boxes], cls.mode))

    def __getitem__(self, idxs):
        return OracleRel(F.getitem(self.tensor, idxs))

    def __add__(self, other):
        return OracleRel(F.union(self.tensor, other.tensor))

    def __bool__(self):
        if self.tensor.shape:
            raise ValueError("Box is not a scalar.")
        return F.as_scalar(self.tensor)

    def to_set(self):
        return frozenset(tuple(int(i) for i in p) for p in np.argwhere(self.tensor))

    def __eq__(self, other):
        return np.array_equal(self.tensor, other.tensor)

    def __hash__(self):
        return xxhash.xxh64(self.tensor.data).intdigest()
