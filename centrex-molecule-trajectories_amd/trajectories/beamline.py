"""Beamline: a z0-sorted collection of elements (beamline.py:10-90 of the reference)."""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import List

from scipy.constants import g as _G_SCIPY


@dataclass
class Beamline:
    """Elements are sorted by z0 at construction (stable sort, beamline.py:17-18,44-45).

    ``g`` (keyword extension) is the gravitational acceleration along -Y; the reference hard-codes
    scipy.constants.g."""

    elements: List
    g: float = _G_SCIPY

    def __post_init__(self):
        self.sort_elements()

    def sort_elements(self):
        """Sort the elements in the order molecules fly through them (stable in z0)."""
        self.elements.sort(key=lambda e: e.z0)

    def propagate_through(self, molecule) -> None:
        """Propagate one molecule through every element on the device (beamline.py:20-38): the
        molecule's trajectory gains every row the reference would store, its label is the stopping
        element or "Detected", and NaN rows are dropped."""
        from .molecule import propagate_molecules_through

        propagate_molecules_through(self.elements, [molecule], g=self.g, set_detected=True)
        molecule.trajectory.drop_nans()

    def find_element(self, name):
        for element in self.elements:
            if element.name == name:
                return element
        print(f"Element with name '{name}' not found in beamline")

    def plot(self):
        import matplotlib.pyplot as plt

        fig, axes = plt.subplots(2, 1, figsize=(16, 9))
        axes[0].set_ylabel("X-position / m")
        axes[0].set_ylim([-0.06, 0.06])
        axes[1].set_xlabel("Z-position / m")
        axes[1].set_ylabel("Y-position / m")
        axes[1].set_ylim([-0.06, 0.06])
        for ax in axes:
            ax.set_xlim([0, self.elements[-1].z1 + 0.1])
        for element in self.elements:
            element.plot(axes)
        return axes

    def save_to_hdf(self, filepath: Path, run_name: str) -> None:
        from .hdf import save_beamline

        save_beamline(self, filepath, run_name)
