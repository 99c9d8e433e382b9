"""
Drop-in name for code written against ``tensorsat.contractions.cotengra``
(``tensorsat/contractions/cotengra.py:52``): ``CotengraContraction`` is this package's
:class:`~tensorsat_amd.contractions.planned.PlannedContraction`; the tree comes from the
package's own planner, not from cotengra.
"""

from .planned import PlannedContraction as CotengraContraction

__all__ = ("CotengraContraction",)
