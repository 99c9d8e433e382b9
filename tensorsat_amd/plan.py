"""
Contraction planning: Python side of ``libtsplan.so`` (``include/tensorsat_plan.h``).

Replaces the reference's cotengra planner call ``ctg.array_contract_tree``
(``contractions/cotengra.py:78-85``) and ``ContractionTree.slice_and_reconfigure``
(``notebooks/02-random-ksat.ipynb`` cell 32).  A :class:`Plan` holds an SSA path
(leaves ``0..S-1``, the ``i``-th merge creates ``S+i``), the sliced wires, and the
cotengra-style cost statistics.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .network import Wiring

_LIB_PATH = Path(__file__).resolve().parent / "libtsplan.so"
_lib = None


class _Net(ctypes.Structure):
    _fields_ = [
        ("num_slots", ctypes.c_int32),
        ("slot_offsets", ctypes.POINTER(ctypes.c_int32)),
        ("slot_wires", ctypes.POINTER(ctypes.c_int32)),
        ("num_wires", ctypes.c_int32),
        ("wire_dims", ctypes.POINTER(ctypes.c_int32)),
        ("num_out", ctypes.c_int32),
        ("out_wires", ctypes.POINTER(ctypes.c_int32)),
    ]


class _Cost(ctypes.Structure):
    _fields_ = [
        ("log2_flops", ctypes.c_double),
        ("log2_write", ctypes.c_double),
        ("log2_max_size", ctypes.c_double),
        ("log2_peak", ctypes.c_double),
        ("num_slices", ctypes.c_double),
    ]


class _Hyper(ctypes.Structure):
    _fields_ = [
        ("trials", ctypes.c_int32),
        ("costmod_min", ctypes.c_double),
        ("costmod_max", ctypes.c_double),
        ("temperature_min", ctypes.c_double),
        ("temperature_max", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("threads", ctypes.c_int32),
        ("write_weight", ctypes.c_double),
        ("log2_size_limit", ctypes.c_double),
    ]


def lib():
    """Load ``libtsplan.so``; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(
                f"{_LIB_PATH} is missing: build it with `python -m tensorsat_amd.build plan`"
            )
        L = ctypes.CDLL(str(_LIB_PATH))
        P = ctypes.POINTER
        L.tsp_greedy.argtypes = [P(_Net), ctypes.c_double, ctypes.c_double, ctypes.c_uint64, P(ctypes.c_int32)]
        L.tsp_hyper.argtypes = [P(_Net), P(_Hyper), P(ctypes.c_int32), P(_Cost)]
        L.tsp_tree_cost.argtypes = [P(_Net), P(ctypes.c_int32), ctypes.c_int32, P(ctypes.c_int32), P(_Cost)]
        L.tsp_slice.argtypes = [P(_Net), P(ctypes.c_int32), ctypes.c_double, ctypes.c_int32,
                                P(ctypes.c_int32), P(ctypes.c_int32), P(_Cost)]
        for f in (L.tsp_greedy, L.tsp_hyper, L.tsp_tree_cost, L.tsp_slice):
            f.restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc == -1:
        raise ValueError(f"{what}: invalid network or path")
    if rc == -2:
        raise MemoryError(f"{what}: out of memory")
    if rc != 0:
        raise RuntimeError(f"{what}: error {rc}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


class _NetHolder:
    """Keeps the int32 arrays alive while the C struct points at them."""

    def __init__(self, wiring: Wiring, drop_wires=()):
        offs, flat, outw, dims = wiring.flat_arrays()
        if drop_wires:
            keep = ~np.isin(flat, np.asarray(list(drop_wires), dtype=np.int32))
            counts = np.diff(offs)
            seg = np.repeat(np.arange(len(counts)), counts)
            new_counts = np.bincount(seg[keep], minlength=len(counts)).astype(np.int32)
            flat = np.ascontiguousarray(flat[keep])
            offs = np.zeros_like(offs)
            offs[1:] = np.cumsum(new_counts)
        self.arrays = (np.ascontiguousarray(offs), np.ascontiguousarray(flat),
                       np.ascontiguousarray(outw), np.ascontiguousarray(dims))
        o, f, ow, d = self.arrays
        self.net = _Net(wiring.num_slots, _ptr(o), _ptr(f), wiring.num_wires, _ptr(d), len(ow), _ptr(ow))


@dataclass
class Cost:
    log2_flops: float
    log2_write: float
    log2_max_size: float
    log2_peak: float
    num_slices: float = 1.0

    @classmethod
    def _from(cls, c: _Cost) -> "Cost":
        return cls(c.log2_flops, c.log2_write, c.log2_max_size, c.log2_peak, c.num_slices)


@dataclass
class Plan:
    """An SSA contraction path over a wiring, with optional sliced wires."""

    num_slots: int
    path: np.ndarray  # int32 [S-1, 2]
    sliced: tuple[int, ...] = ()
    cost: Cost | None = None
    meta: dict = field(default_factory=dict)

    def linear_path(self) -> list[tuple[int, int]]:
        """The same tree as an opt_einsum-style pop/append path
        (``contractions/simple.py:41-50``)."""
        live = list(range(self.num_slots))
        out = []
        for k, (a, b) in enumerate(self.path.tolist()):
            i, j = live.index(a), live.index(b)
            out.append((i, j))
            for x in sorted((i, j), reverse=True):
                live.pop(x)
            live.append(self.num_slots + k)
        return out

    @classmethod
    def from_linear_path(cls, num_slots: int, lpath) -> "Plan":
        live = list(range(num_slots))
        ssa = []
        for k, (i, j) in enumerate(lpath):
            a = live[i]
            b = live[j]
            ssa.append((a, b))
            for x in sorted((i, j), reverse=True):
                live.pop(x)
            live.append(num_slots + k)
        return cls(num_slots, np.asarray(ssa, dtype=np.int32).reshape(-1, 2))


def greedy(wiring: Wiring, *, costmod: float = 1.0, temperature: float = 0.0, seed: int = 0) -> Plan:
    """One greedy run (cotengra ``optimize="greedy"`` counterpart)."""
    h = _NetHolder(wiring)
    S = wiring.num_slots
    path = np.zeros(max(0, 2 * (S - 1)), dtype=np.int32)
    _check(lib().tsp_greedy(ctypes.byref(h.net), costmod, temperature, seed, _ptr(path)), "tsp_greedy")
    plan = Plan(S, path.reshape(-1, 2))
    plan.cost = tree_cost(wiring, plan)
    return plan


def hyper(
    wiring: Wiring,
    *,
    trials: int = 64,
    seed: int = 0,
    threads: int = 0,
    costmod: tuple[float, float] = (0.1, 4.0),
    temperature: tuple[float, float] = (0.001, 1.0),
    write_weight: float = 64.0,
    log2_size_limit: float = math.inf,
    drop_wires=(),
) -> Plan:
    """Hyper-optimised random greedy (cotengra ``optimize="auto"``/hyper counterpart)."""
    h = _NetHolder(wiring, drop_wires)
    S = wiring.num_slots
    path = np.zeros(max(0, 2 * (S - 1)), dtype=np.int32)
    opts = _Hyper(trials, costmod[0], costmod[1], temperature[0], temperature[1], seed, threads,
                  write_weight, log2_size_limit)
    c = _Cost()
    _check(lib().tsp_hyper(ctypes.byref(h.net), ctypes.byref(opts), _ptr(path), ctypes.byref(c)), "tsp_hyper")
    return Plan(S, path.reshape(-1, 2), cost=Cost._from(c))


def tree_cost(wiring: Wiring, plan: Plan, sliced=None) -> Cost:
    h = _NetHolder(wiring)
    sl = np.asarray(plan.sliced if sliced is None else sliced, dtype=np.int32)
    path = np.ascontiguousarray(plan.path.reshape(-1), dtype=np.int32)
    c = _Cost()
    _check(lib().tsp_tree_cost(ctypes.byref(h.net), _ptr(path), len(sl), _ptr(sl), ctypes.byref(c)), "tsp_tree_cost")
    return Cost._from(c)


def slice_plan(wiring: Wiring, plan: Plan, target_log2_size: float, *, max_sliced: int = 64,
               reconfigure_trials: int = 0, seed: int = 0) -> Plan:
    """
    Choose sliced wires so the largest intermediate is <= 2^target, then (when
    ``reconfigure_trials`` > 0) re-plan the tree with those wires fixed and keep the
    cheaper of the two trees that meets the target.
    """
    h = _NetHolder(wiring)
    path = np.ascontiguousarray(plan.path.reshape(-1), dtype=np.int32)
    out = np.zeros(max_sliced, dtype=np.int32)
    n = ctypes.c_int32(0)
    c = _Cost()
    _check(lib().tsp_slice(ctypes.byref(h.net), _ptr(path), target_log2_size, max_sliced, _ptr(out),
                           ctypes.byref(n), ctypes.byref(c)), "tsp_slice")
    sliced = tuple(int(w) for w in out[: n.value])
    best = Plan(plan.num_slots, plan.path.copy(), sliced, Cost._from(c), dict(plan.meta))
    if reconfigure_trials > 0 and sliced:
        alt = hyper(wiring, trials=reconfigure_trials, seed=seed, drop_wires=sliced,
                    log2_size_limit=target_log2_size)
        alt.sliced = sliced
        alt.cost = tree_cost(wiring, alt)
        if alt.cost.log2_max_size <= target_log2_size + 1e-9 and (
            best.cost.log2_max_size > target_log2_size + 1e-9 or alt.cost.log2_flops < best.cost.log2_flops
        ):
            best = alt
    return best
