"""Beamline elements: the BeamlineElement interface and the apertures.

Same dataclass fields and methods as the reference (apertures.py:22-286).  ``propagate_through``
runs the element on the device (a one-element beamline launched from the molecule's current row);
``_ctraj_pack`` emits the element's row of the C-ABI element table.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass
from pathlib import Path

from .. import _native as N

__all__ = ["BeamlineElement", "CircularAperture", "RectangularAperture", "FieldPlates"]


@dataclass
class BeamlineElement(ABC):
    """Interface of every element (apertures.py:22-80)."""

    name: str  # element name; it is the Counter label of molecules stopped here
    z0: float  # start of the element along Z (m)
    L: float  # length along Z (m)
    x0: float = 0.0  # centre X (m)
    y0: float = 0.0  # centre Y (m)

    def __post_init__(self):
        self.z1 = self.z0 + self.L

    def propagate_through(self, molecule) -> None:
        """Propagate one molecule through this element on the device (apertures.py:38-42)."""
        from ..molecule import propagate_molecules_through

        propagate_molecules_through([self], [molecule], g=getattr(molecule, "g", None), set_detected=False)

    @abstractmethod
    def N_steps(self) -> int:
        """Maximum trajectory rows this element appends (apertures.py:44-48)."""

    @abstractmethod
    def plot(self, axes):
        """Draw the element (axes[0]: XZ, axes[1]: YZ)."""

    @abstractmethod
    def _ctraj_pack(self, reg) -> N.Elem:
        """Row of the device element table."""

    def save_to_hdf(self, filepath: Path, parent_group_path: str) -> None:
        from ..hdf import save_element

        save_element(self, filepath, parent_group_path)


def _rect(axes_i, xy, w, h, color):
    from matplotlib.patches import Rectangle

    axes_i.add_patch(Rectangle(xy, w, h, color=color))


@dataclass
class CircularAperture(BeamlineElement):
    """Circular hole of diameter d, tested at z0 and z1: a molecule stops iff sqrt(x^2+y^2) > d/2
    (apertures.py:92-115; the centre x0, y0 is not used by that test)."""

    d: float = 0.0254  # diameter (m)

    def N_steps(self):
        return 2

    def plot(self, axes):
        grey = (0.5, 0.5, 0.5)
        for ax in axes[:2]:
            _rect(ax, (self.z0, self.d / 2), self.z1 - self.z0, 1, grey)
            _rect(ax, (self.z0, -self.d / 2 - 1), self.z1 - self.z0, 1, grey)

    def _ctraj_pack(self, reg):
        E = N.Elem()
        E.kind = N.CIRC
        E.label_hit = reg.label(self.name)
        E.z0, E.z1 = float(self.z0), float(self.z1)
        E.p[0] = self.d / 2
        return E


@dataclass
class RectangularAperture(BeamlineElement):
    """Box x1 < x < x2, y1 < y < y2 (strict), tested at z0 and z1 (apertures.py:165-189).
    As in the reference, ``w`` is the X extent and ``h`` the Y extent."""

    w: float = 0.02
    h: float = 0.02

    def __post_init__(self):
        super().__post_init__()
        self.x1 = self.x0 - self.w / 2
        self.x2 = self.x0 + self.w / 2
        self.y1 = self.y0 - self.h / 2
        self.y2 = self.y0 + self.h / 2

    def N_steps(self):
        return 2

    def plot(self, axes):
        _rect(axes[0], (self.z0, self.x2), self.z1 - self.z0, 0.05, "k")
        _rect(axes[0], (self.z0, self.x1 - 0.05), self.z1 - self.z0, 0.05, "k")
        _rect(axes[1], (self.z0, self.y2), self.z1 - self.z0, 0.05, "k")
        _rect(axes[1], (self.z0, self.y1 - 0.05), self.z1 - self.z0, 0.05, "k")

    def _ctraj_pack(self, reg):
        E = N.Elem()
        E.kind = N.RECT
        E.label_hit = reg.label(self.name)
        E.z0, E.z1 = float(self.z0), float(self.z1)
        E.p[0], E.p[1], E.p[2], E.p[3] = self.x1, self.x2, self.y1, self.y2
        return E


@dataclass
class FieldPlates(BeamlineElement):
    """Plates at x1 = x0 - w/2 and x2 = x0 + w/2 from z0 to z1; a molecule that reaches a plate stops
    at the plate (apertures.py:227-270)."""

    w: float = 0.02

    def __post_init__(self):
        super().__post_init__()
        self.x1 = self.x0 - self.w / 2
        self.x2 = self.x0 + self.w / 2

    def N_steps(self):
        return 2

    def plot(self, axes):
        _rect(axes[0], (self.z0, self.x2), self.z1 - self.z0, 0.02, "y")
        _rect(axes[0], (self.z0, self.x1 - 0.02), self.z1 - self.z0, 0.02, "y")

    def _ctraj_pack(self, reg):
        E = N.Elem()
        E.kind = N.PLATES
        E.label_hit = reg.label(self.name)
        E.z0, E.z1 = float(self.z0), float(self.z1)
        E.p[0], E.p[1] = self.x1, self.x2
        return E
