"""Contractions: the ``Contraction`` protocol (``tensorsat/contractions/abc.py:29-97``),
explicit-path contractions and planned (tree) contractions."""

from .abc import Contraction
from .planned import PlannedContraction
from .simple import SimpleContraction

__all__ = ("Contraction", "PlannedContraction", "SimpleContraction")
