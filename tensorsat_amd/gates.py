"""
Leaf tensors of SAT networks: the binary gates and constants of the reference's
``tensorsat/lib/bincirc.py:28-72``, as read-only uint8 0/1 arrays.

Axis convention follows ``FinRel.from_callable`` (``lang/fin_rel.py:228-249``):
input axes first, output axis last; entry 1 where ``out == f(inputs)``.
"""

from __future__ import annotations

import itertools

import numpy as np


def _from_callable(n_in: int, n_out: int, func) -> np.ndarray:
    t = np.zeros((2,) * (n_in + n_out), dtype=np.uint8)
    for idx in itertools.product(range(2), repeat=n_in):
        out = func(*idx)
        if not isinstance(out, tuple):
            out = (out,)
        t[idx + out] = 1
    t.setflags(write=False)
    return t


def _from_set(points, rank: int) -> np.ndarray:
    t = np.zeros((2,) * rank, dtype=np.uint8)
    for p in points:
        t[p] = 1
    t.setflags(write=False)
    return t


not_ = _from_callable(1, 1, lambda a: 1 - a)
and_ = _from_callable(2, 1, lambda a, b: a & b)
or_ = _from_callable(2, 1, lambda a, b: a | b)
impl_ = _from_callable(2, 1, lambda a, b: (1 - a) | b)
biimpl_ = _from_callable(2, 1, lambda a, b: (a & b) | ((1 - a) & (1 - b)))
xor_ = _from_callable(2, 1, lambda a, b: a ^ b)
bit_0 = _from_set([(0,)], 1)
bit_1 = _from_set([(1,)], 1)
bit_unk = _from_set([(0,), (1,)], 1)

GATES = {
    "not_": not_,
    "and_": and_,
    "or_": or_,
    "impl_": impl_,
    "biimpl_": biimpl_,
    "xor_": xor_,
    "bit_0": bit_0,
    "bit_1": bit_1,
    "bit_unk": bit_unk,
}


def clause_tensor(signs: tuple[bool, ...]) -> np.ndarray:
    """
    The (2,)*k clause tensor over distinct variables: all ones except a single 0 at
    the falsifying assignment (variable = 0 for a positive literal, 1 for a negative
    one).  ``signs[i]`` is True for a positive literal.
    """
    k = len(signs)
    t = np.ones((2,) * k, dtype=np.uint8)
    t[tuple(0 if s else 1 for s in signs)] = 0
    return t
