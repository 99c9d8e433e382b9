"""
Satisfying assignments (witnesses) of CNF instances.

The reference finds an assignment by self-reduction over one contraction: the wiring of
``inputs(x) >> diagram()`` is the same for every input selection x
(``CNFInstance.contraction_wiring``, ``lib/sat.py:231-246``), so one planned contraction
decides, variable by variable in index order, whether the instance stays satisfiable
with that variable 0 (else it is 1) -- up to 2n contractions with different ``bit_0`` /
``bit_1`` / ``bit_unk`` leaves (``notebooks/02-random-ksat.ipynb`` cell 36,
``01-introduction.ipynb`` cell 113).  The result is the lexicographically smallest
satisfying assignment.

* :func:`find_assignment` is that loop, on any box class (``DeviceRel`` runs every
  contraction on the GPU).
* :func:`find_assignments` returns the same assignments for a batch at once: the device
  cube search branches on the smallest residual variable, false first, one job per
  instance, and its first satisfiable leaf -- the assignment forced on the way down, the
  lexicographically smallest entry of the leaf's relation, 0 for variables left free --
  is that assignment.
"""

from __future__ import annotations

from .sat import CNFInstance


def find_assignment(cnf: CNFInstance, *, box_class=None, optimize: str = "auto") -> str | None:
    """The lexicographically smallest satisfying assignment ("0"/"1" per variable), or
    None if the instance is unsatisfiable, by the reference's 2n-contraction loop."""
    from .contractions import PlannedContraction

    if box_class is None:
        from .rel import DeviceRel as box_class
    n = cnf.num_vars
    cd = cnf.diagram()
    contraction = PlannedContraction(box_class, (cnf.inputs(None) >> cd).flatten().wiring, optimize=optimize)

    def sat(spec: str) -> bool:
        return bool(contraction.contract((cnf.inputs(spec) >> cd).flatten()))

    if not sat("?" * n):
        return None
    assignment = ""
    for i in range(n):
        rest = "?" * (n - i - 1)
        assignment += "0" if sat(assignment + "0" + rest) else "1"
    return assignment


def find_assignments(cnfs, *, width: int = 6) -> list[str | None]:
    """:func:`find_assignment` for a batch, by the device cube search (module docstring)."""
    from .search import SearchSolver

    s = SearchSolver(width=width, split_target=None, budget=None, order="lex", witness=True)
    s.run(cnfs)
    return s.witnesses()
