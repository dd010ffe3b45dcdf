"""Electrostatic lens element (electrostatic_lens.py:23-228).

Fields d, dz, V, a_interp, state, mass mean what they mean in the reference.  Keyword extensions
select the configuration-mode physics the reference does not have:

* ``stark``      -- a :class:`~trajectories.stark.WSpline` (or a list, one per molecular state): the
                    force is -grad W(|E|)/m with W' taken from the spline;
* ``field``      -- "linear" (|E| = 2 V r / (d/2)^2, the reference's assumption, :199-200) or
                    "hexapole" (|E| = 3 V r^2 / (d/2)^3);
* ``integrator`` -- "ref" (N = rint(L/dz) steps of dz/vz with the reference's RK4 variant, :86-111)
                    or "rk4" (classical RK4 at fixed ``dt``, then one partial step to z1).

Without ``stark`` the force is the reference's radial-acceleration table: ``a_interp`` if given
(a linear ``scipy.interpolate.interp1d`` or any object with ``x`` / ``y`` knot arrays), otherwise
built exactly as the reference builds it (:193-206) from ``state.stark_potential(E_V_per_cm)``.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

from .. import _native as N
from ..compile import LABEL_INSIDE_LENS, LABEL_LENS_ENTRANCE
from .apertures import BeamlineElement, _rect

__all__ = ["ElectrostaticLens", "LinearTable"]


class LinearTable:
    """Knots (x, y) evaluated by linear interpolation (np.interp), like the reference's interp1d."""

    def __init__(self, x, y):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)

    def __call__(self, r):
        r = np.asarray(r, dtype=np.float64)
        if np.any(r < self.x[0]) or np.any(r > self.x[-1]):
            raise ValueError("A value in x_new is outside the interpolation range.")
        return np.interp(r, self.x, self.y)


@dataclass
class ElectrostaticLens(BeamlineElement):
    d: float = 1.75 * 0.0254  # bore diameter (m)
    dz: float = 1e-3  # spatial integration step of the "ref" integrator (m)
    V: float = 27.6e3  # electrode voltage (V)
    a_interp: Any = None  # radial acceleration table a_r(r)
    state: Any = None  # molecular state; needs .stark_potential(E_V_per_cm) -> J to build a_interp
    mass: float = (204.38 + 19.00) * 1.67e-27  # kg
    stark: Any = None
    field: str = "linear"
    integrator: str = "ref"
    dt: float = 10e-6  # time step of the "rk4" integrator (s)

    def N_steps(self):
        """Rows appended: 1 entrance row plus one per integration step (electrostatic_lens.py:120-125).
        For the "rk4" integrator the count depends on v_z; this is the count at v_z = dz/dt."""
        if self.integrator == "ref":
            return 1 + int(np.rint(self.L / self.dz))
        return 2 + int(np.floor(self.L / self.dz))

    def rows_needed(self, vz_min: float) -> int:
        """Upper bound on the rows this lens appends for molecules with v_z >= vz_min."""
        if self.integrator == "ref":
            return self.N_steps()
        return 3 + int(np.floor(self.L / (vz_min * self.dt)))

    def plot(self, axes):
        for ax in axes[:2]:
            _rect(ax, (self.z0, self.d / 2), self.z1 - self.z0, 0.02, "b")
            _rect(ax, (self.z0, -self.d / 2 - 0.02), self.z1 - self.z0, 0.02, "b")

    def save_to_hdf(self, filepath: Path, parent_group_path: str) -> None:
        from ..hdf import save_element

        if self.stark is None and self.a_interp is None and hasattr(self.state, "stark_potential"):
            self.radial_table()  # the state is stored as its repr: keep the table it defines
        save_element(self, filepath, parent_group_path, skip=("a_interp", "stark"), repr_keys=("state",))

    # ---------------------------------------------------------------- force tables
    def field_coefficient(self) -> float:
        """cE of |E| = cE r^p (p = 1 linear, 2 hexapole), from |V|."""
        R = self.d / 2
        if self.field == "hexapole":
            return 3.0 * abs(self.V) / R ** 3
        if self.field == "linear":
            return 2.0 * abs(self.V) / R ** 2
        raise ValueError(f"unknown field model {self.field!r}")

    def radial_table(self):
        """(r, a_r) knots of the radial-acceleration table."""
        if self.a_interp is None:
            self.a_interp = self._build_radial_table()
        t = self.a_interp
        kind = getattr(t, "_kind", getattr(t, "kind", "linear"))
        if kind != "linear" or not hasattr(t, "x") or not hasattr(t, "y"):
            raise TypeError("a_interp must be a linear interp1d (or have x / y knot arrays) to run on the device")
        y = np.asarray(t.y, dtype=np.float64)
        if y.ndim != 1:
            raise TypeError("a_interp must be 1-D")
        return np.asarray(t.x, dtype=np.float64), y

    def _build_radial_table(self):
        # the construction of electrostatic_lens.py:193-209 with a pluggable Stark potential
        if self.state is None or not hasattr(self.state, "stark_potential"):
            raise NotImplementedError(
                "ElectrostaticLens needs a_interp, a stark spline, or a state with .stark_potential(E_V_per_cm); "
                "the centrex_TlF Hamiltonian diagonalisation is not part of this engine")
        dr = 1e-4
        r_values = np.linspace(0, self.d / 2 * 1.01, int(np.round(self.d / 2 / dr)))
        if self.field == "hexapole":
            E_values = 3 * self.V / ((self.d / 2) ** 3) * r_values ** 2
        else:
            E_values = 2 * self.V / ((self.d / 2) ** 2) * r_values
        V_stark = self.state.stark_potential(E_values / 100)
        a_values = -np.gradient(V_stark, dr) / self.mass
        return LinearTable(r_values, a_values)

    def lens_acceleration(self, x):
        """Acceleration at position x (host evaluation of the same force the device uses)."""
        x = np.asarray(x, dtype=np.float64)
        g = 9.80665
        r = float(np.sqrt(np.sum(x[:2] ** 2)))
        a = np.zeros(3)
        if self.stark is None:
            rr, aa = self.radial_table()
            ar = np.interp(r, rr, aa)
            if r != 0:
                a[0] = ar * x[0] / r
                a[1] = ar * x[1] / r
        else:
            spl = self.stark[0] if isinstance(self.stark, (list, tuple)) else self.stark
            cE = self.field_coefficient()
            if self.field == "hexapole":
                f = -spl.dW(cE * r * r) * 2 * cE / self.mass
                a[0], a[1] = f * x[0], f * x[1]
            elif r != 0:
                f = -spl.dW(cE * r) * cE / (self.mass * r)
                a[0], a[1] = f * x[0], f * x[1]
        a[1] -= g
        return a

    def _ctraj_pack(self, reg):
        E = N.Elem()
        E.kind = N.LENS
        E.label_hit = reg.label(LABEL_LENS_ENTRANCE)
        E.label_inside = reg.label(LABEL_INSIDE_LENS)
        E.z0, E.z1 = float(self.z0), float(self.z1)
        R = self.d / 2
        E.p[0] = R
        E.p[1] = R * R
        if self.integrator == "ref":
            E.integrator = N.INTEG_REF
            E.n_steps = int(np.rint(self.L / self.dz))
            E.p[2] = float(self.dz)
        elif self.integrator == "rk4":
            E.integrator = N.INTEG_RK4
            if not self.dt > 0:
                raise ValueError("dt must be > 0")
            E.p[2] = float(self.dt)
        else:
            raise ValueError(f"unknown integrator {self.integrator!r}")
        E.p[3] = float(self.L)
        E.p[7] = float(self.V)
        if self.stark is None:
            r, a = self.radial_table()
            E.force = N.FORCE_RTABLE
            E.table = reg.rtable(r, a)
        else:
            reg.wspline(self.stark)
            spl = reg.wsplines[0]
            cE = self.field_coefficient()
            if self.field == "hexapole":
                E.force = N.FORCE_WSPLINE_HEX
                E.p[4] = -2.0 * cE / self.mass
            else:
                E.force = N.FORCE_WSPLINE_LIN
                E.p[4] = -cE / self.mass
            E.p[5] = cE / spl.dE
        return E
