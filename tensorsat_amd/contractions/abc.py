"""
Abstract contraction (``tensorsat/contractions/abc.py:29-97``): a contraction is built
for a box class and a wiring, and contracts any flat diagram with that wiring.
"""

from __future__ import annotations

from abc import ABC, abstractmethod

from ..network import Diagram, Wiring


class Contraction(ABC):
    """Abstract base class for contractions."""

    def __init__(self, box_class: type, wiring: Wiring) -> None:
        if not isinstance(wiring, Wiring):
            raise TypeError("wiring must be a Wiring")
        if not getattr(box_class, "can_be_contracted", lambda: False)():
            raise ValueError("Given box class cannot be contracted.")
        self._box_class = box_class
        self._wiring = wiring

    @property
    def box_class(self) -> type:
        return self._box_class

    @property
    def wiring(self) -> Wiring:
        return self._wiring

    def can_contract(self, diagram: Diagram) -> bool:
        try:
            self.validate(diagram)
            return True
        except ValueError:
            return False

    def validate(self, diagram: Diagram) -> None:
        """Same conditions as ``abc.py:64-89``: matching wiring, no open slots, flat."""
        if not isinstance(diagram, Diagram):
            raise TypeError("diagram must be a Diagram")
        if diagram.wiring != self.wiring:
            raise ValueError("Diagram's wiring must match contraction wiring.")
        if diagram.num_open_slots > 0:
            raise ValueError("Diagram cannot have open slots.")
        if not diagram.is_flat:
            raise ValueError("Diagram must be flat.")

    def _boxes(self, diagram: Diagram) -> list:
        """The diagram's blocks as boxes of this contraction's class (host arrays are
        uploaded once)."""
        cls = self._box_class
        out = []
        for b in diagram.blocks:
            if isinstance(b, cls):
                out.append(b)
            else:
                conv = getattr(cls, "from_numpy", None)
                if conv is None:
                    raise ValueError(
                        f"Cannot contract diagram: block of type {type(b).__name__} is not a {cls.__name__}"
                    )
                out.append(conv(b))
        return out

    @abstractmethod
    def contract(self, diagram: Diagram, *args, **kwargs):
        """Validate and contract the diagram."""
