"""
CNF instances: random k-SAT generation, DIMACS input/output and input selections.

Mirrors ``CNFInstance`` of the reference (``tensorsat/lib/sat.py:40-275``): the same
constructor names, argument meaning, validation errors and random stream, so that a
seed produces the same clauses as the reference generator.  Diagram building is done
by :mod:`tensorsat_amd.network`, which emits flat index lists directly instead of a
Python object graph.
"""

from __future__ import annotations

import re
from collections.abc import Sequence
from math import comb

import numpy as np

Clause = tuple[int, ...]
"""A clause: non-zero literals, sign = polarity, magnitude = 1-based variable."""


class CNFInstance:
    """A SAT instance in conjunctive normal form (``lib/sat.py:40``)."""

    __slots__ = ("_num_vars", "_clauses", "_arrays", "_addr", "_head", "_hb")

    def __init__(self, num_vars: int, clauses: Sequence[Sequence[int]]) -> None:
        """Public constructor, validating literals (``lib/sat.py:111-121``)."""
        clauses = tuple(tuple(int(l) for l in c) for c in clauses)
        lits = [abs(l) for c in clauses for l in c]
        if lits and num_vars < max(lits):
            raise ValueError("Clauses contain invalid variables.")
        if any(l == 0 for c in clauses for l in c):
            raise ValueError("Literals must be non-zero.")
        self._num_vars = int(num_vars)
        self._clauses = clauses
        self._arrays = None
        self._addr = None
        self._head = None
        self._hb = None

    @classmethod
    def _new(cls, num_vars: int, clauses: tuple[Clause, ...], arrays=None) -> "CNFInstance":
        self = object.__new__(cls)
        self._num_vars = num_vars
        self._clauses = clauses
        self._arrays = arrays
        self._addr = None
        self._head = None
        self._hb = None
        return self

    @classmethod
    def random(
        cls, k: int, n: int, m: int, *, rng: int | np.random.Generator | None = None
    ) -> "CNFInstance":
        """
        Random k-SAT with ``n`` variables and ``m`` distinct clauses.

        Follows ``lib/sat.py:43-78``: per clause, ``k`` distinct variables drawn by
        ``rng.choice(range(n), size=k, replace=False)``, then ``k`` signs drawn by
        ``rng.choice(range(2), size=k)``; literal ``v+1`` for sign 0 and ``-(v+1)``
        for sign 1; literals sorted; a clause equal to an earlier one is redrawn.
        The two ``choice`` calls are made in the same order with the same arguments,
        so a given seed yields the same clause sequence.
        """
        if k <= 0:
            raise ValueError("Clause size 'k' must be positive.")
        if n < k:
            raise ValueError(f"Number of variables 'n' must be at least {k = }.")
        if m <= 0:
            raise ValueError("Number of clauses 'm' must be positive.")
        if m > comb(n, k) * 2**k:
            raise ValueError(f"Number of clauses 'm' cannot exceed {comb(n, k)*2**k = }")
        if not isinstance(rng, np.random.Generator):
            rng = np.random.default_rng(rng)
        clauses: list[Clause] = []
        seen: set[Clause] = set()
        pool = range(n)
        while len(clauses) < m:
            vs = rng.choice(pool, size=k, replace=False)
            signs = rng.choice(range(2), size=k)
            clause = tuple(
                sorted(int(v) + 1 if int(s) == 0 else -int(v) - 1 for v, s in zip(vs, signs))
            )
            if clause not in seen:
                seen.add(clause)
                clauses.append(clause)
        flat = np.asarray(clauses, dtype=np.int32).reshape(-1)
        return cls._new(n, tuple(clauses), (np.full(m, k, dtype=np.int32), flat))

    @classmethod
    def from_dimacs(cls, dimacs: str) -> "CNFInstance":
        """
        Parses DIMACS text (``lib/sat.py:80-100``): comment lines start with ``c``,
        the header is ``p cnf <vars> <clauses>``, then one clause per line with its
        terminating ``0`` dropped.  Literals are separated by single spaces, as the
        reference reads them: other separators (tabs, runs of spaces inside a line)
        raise ``ValueError``, and so does an instance without literals.
        """
        lines = [
            s for line in dimacs.strip().split("\n") if (s := line.strip()) and not s.startswith("c")
        ]
        if not lines:
            raise ValueError("DIMACS code must start with 'p cnf <num vars> <num clauses>'.")
        match = re.compile(r"p cnf ([0-9]+) ([0-9]+)").match(lines[0])
        if not match:
            raise ValueError("DIMACS code must start with 'p cnf <num vars> <num clauses>'.")
        num_vars, num_clauses = map(int, match.groups())
        clauses = tuple(tuple(map(int, line.split(" ")))[:-1] for line in lines[1:])
        if len(clauses) != num_clauses:
            raise ValueError("Number of clauses does not match the specified number.")
        lits = [abs(l) for c in clauses for l in c]
        if not lits:
            raise ValueError("DIMACS code holds no literals.")
        if num_vars < max(lits):
            raise ValueError("Clauses contain invalid variables.")
        return cls._new(num_vars, clauses)

    @property
    def num_vars(self) -> int:
        return self._num_vars

    @property
    def clauses(self) -> tuple[Clause, ...]:
        return self._clauses

    def to_dimacs(self) -> str:
        """DIMACS text (``lib/sat.py:135-142``)."""
        lines = [f"p cnf {self.num_vars} {len(self.clauses)}"]
        lines += [f"{' '.join(map(str, c))} 0" for c in self.clauses]
        return "\n".join(lines)

    def normalize_inputs(self, inputs: str | int | None) -> str:
        """
        Input selection string over ``0``, ``1``, ``?`` (unknown) and ``_`` (open),
        as ``lib/sat.py:144-163``: an int is written in binary, zero-padded to ``n``.
        """
        n = self.num_vars
        if isinstance(inputs, (int, np.integer)) and not isinstance(inputs, bool):
            if inputs < 0:
                raise ValueError("Inputs must be specified by a non-negative number.")
            inputs = f"{int(inputs):0>{n}b}"
        elif inputs is None:
            inputs = "?" * n
        else:
            if not isinstance(inputs, str):
                raise TypeError("Inputs must be a string, an int or None.")
            if not all(b in "01?_" for b in inputs):
                raise ValueError("Input specification characters must be '0', '1', '?' or '_'.")
            if len(inputs) != n:
                raise ValueError(
                    f"Expected input specification string to be {n} characters long,"
                    f" found {len(inputs)} characters instead."
                )
        return inputs

    def literal_arrays(self) -> tuple[np.ndarray, np.ndarray]:
        """(clause lengths, concatenated literals) as int32 arrays, computed once (the
        random generator fills them as it draws), for batch packing without walking
        the clause tuples again."""
        if self._arrays is None:
            lens = np.fromiter((len(c) for c in self._clauses), dtype=np.int32, count=len(self._clauses))
            flat = np.fromiter((l for c in self._clauses for l in c), dtype=np.int32, count=int(lens.sum()))
            self._arrays = (lens, flat)
        return self._arrays

    def literal_addresses(self) -> tuple[int, int, int]:
        """(clause count, address of the lengths, address of the literals) of
        :meth:`literal_arrays`, for native batch packing (valid while the instance
        lives; the arrays are never replaced)."""
        if self._addr is None:
            lens, flat = self.literal_arrays()
            self._addr = (lens.size, lens.ctypes.data, flat.ctypes.data)
        return self._addr

    def pack_header(self) -> tuple[int, int, int, int]:
        """(variable count, clause count, address of the lengths, address of the
        literals): what native batch packing reads per instance."""
        if self._head is None:
            self._head = (self._num_vars, *self.literal_addresses())
        return self._head

    def pack_header_bytes(self) -> bytes:
        """:meth:`pack_header` as four native int64s (cached in ``_hb``, so a batch's
        headers are gathered by one ``bytes.join`` without a Python call per instance)."""
        if self._hb is None:
            self._hb = np.asarray(self.pack_header(), dtype=np.int64).tobytes()
        return self._hb

    def clause_matrix(self) -> np.ndarray:
        """Clauses as a dense ``(m, k_max)`` int32 array, zero-padded on the right."""
        m = len(self.clauses)
        kmax = max((len(c) for c in self.clauses), default=0)
        out = np.zeros((m, kmax), dtype=np.int32)
        for i, c in enumerate(self.clauses):
            out[i, : len(c)] = c
        return out

    def __eq__(self, other: object) -> bool:
        if not isinstance(other, CNFInstance):
            return NotImplemented
        return self._num_vars == other._num_vars and self._clauses == other._clauses

    def __hash__(self) -> int:
        return hash((self._num_vars, self._clauses))

    def __repr__(self) -> str:
        attrs = [f"{self.num_vars} vars", f"{len(self.clauses)} clauses"]
        ks = {len(c) for c in self.clauses}
        if len(ks) == 1:
            attrs.append(f"k={ks.pop()}")
        elif ks:
            attrs.append(f"k in [{min(ks)}..{max(ks)}]")
        return f"<CNFInstance: {', '.join(attrs)}>"


class _Inputs:
    """``cnf.inputs(x)``: an input selection waiting to be composed with ``cnf.diagram()``
    (``lib/sat.py:165-206``; composition ``diagrams.py:1592-1637``)."""

    __slots__ = ("cnf", "spec")

    def __init__(self, cnf: CNFInstance, spec: str) -> None:
        self.cnf = cnf
        self.spec = spec

    def __rshift__(self, other):
        from .network import Diagram, sat_network

        src = getattr(other, "_cnf_source", None) if isinstance(other, Diagram) else None
        if src is None or src[0] != self.cnf:
            raise ValueError("inputs(...) composes with the diagram of the same CNF instance")
        return sat_network(self.cnf, self.spec, mode=src[1])

    @property
    def wiring(self):
        from .network import Wiring

        n = self.cnf.num_vars
        slots = [(i,) for i, b in enumerate(self.spec) if b in "01?"]
        return Wiring((2,) * n, slots, list(range(n)))


def _cnf_diagram(self: CNFInstance, *, mode: str = "bintree"):
    """``CNFInstance.diagram`` (``lib/sat.py:208-229``): every variable an open port."""
    from .network import sat_network

    d = sat_network(self, "_" * self.num_vars, mode=mode)
    object.__setattr__(d, "_cnf_source", (self, mode))
    return d


def _cnf_inputs(self: CNFInstance, inputs: str | int | None, /, *, discard_unk: bool = False) -> _Inputs:
    """``CNFInstance.inputs`` (``lib/sat.py:165-206``); ``discard_unk`` leaves unknown
    variables as plain summed wires instead of ``bit_unk`` leaves."""
    spec = self.normalize_inputs(inputs)
    if discard_unk:
        spec = spec.replace("?", "*")
    return _Inputs(self, spec)


def _contraction_wiring(self: CNFInstance, *, mode: str = "bintree"):
    """``CNFInstance.contraction_wiring`` (``lib/sat.py:231-246``): the wiring shared by
    every ``inputs(x) >> diagram()`` composition with 0/1/? inputs."""
    return (self.inputs(0) >> self.diagram(mode=mode)).wiring


CNFInstance.diagram = _cnf_diagram
CNFInstance.inputs = _cnf_inputs
CNFInstance.contraction_wiring = _contraction_wiring
