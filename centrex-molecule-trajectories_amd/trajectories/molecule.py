"""Molecule and Trajectory (molecule.py:8-185 of the reference).

These are the per-molecule state containers of the reference's API.  Propagating them through
elements happens on the device (:func:`propagate_molecules_through`); the large runs of
``TrajectorySimulator`` never build per-molecule objects except for the molecules it keeps.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.constants import g as _G

from . import _native as N


@dataclass
class Molecule:
    alive: bool = True

    def init_trajectory(self, beamline, x0=np.array((0, 0, 0)), v0=np.array((0, 0, 200)), a0=None, t0=0):
        """Start a trajectory sized for ``beamline`` at (x0, v0, a0, t0) (molecule.py:15-24)."""
        self.g = float(getattr(beamline, "g", _G))
        if a0 is None:
            a0 = np.array((0, -self.g, 0))
        self.trajectory = Trajectory(beamline)
        self.trajectory.update(x0, v0, a0, t0)

    def x(self, delta_t: float = None):
        if not delta_t:
            return self.trajectory.x[self.trajectory.n - 1, :]
        return self.x() + self.v() * delta_t + self.a() * delta_t ** 2 / 2

    def v(self, delta_t: float = None):
        if not delta_t:
            return self.trajectory.v[self.trajectory.n - 1, :]
        return self.v() + self.a() * delta_t

    def a(self):
        return self.trajectory.a[self.trajectory.n - 1, :]

    def t(self):
        return self.trajectory.t[self.trajectory.n - 1]

    def update_trajectory(self, delta_t, a=None):
        """Append the constant-acceleration state after delta_t (molecule.py:58-68)."""
        if a is None:
            a = np.array((0, -getattr(self, "g", _G), 0))
        x = self.x(delta_t)
        v = self.v(delta_t)
        t = self.t() + delta_t
        self.trajectory.update(x, v, a, t)

    def set_aperture_hit(self, aperture_name):
        self.aperture_hit = aperture_name

    def set_dead(self):
        self.alive = False

    def plot_trajectory(self, axes):
        color = "g" if self.aperture_hit == "Detected" else ("r" if self.aperture_hit == "Field plates" else "k")
        axes[0].plot(self.trajectory.x[:, 2], self.trajectory.x[:, 0], c=color)
        axes[1].plot(self.trajectory.x[:, 2], self.trajectory.x[:, 1], c=color)

    def save_to_hdf(self, file, run_name: str, group_name: str):
        self.trajectory.save_to_hdf(file, run_name, group_name)
        file[run_name + "/" + group_name].attrs["aperture_hit"] = self.aperture_hit
        file[run_name + "/" + group_name].attrs["alive"] = self.alive


class Trajectory:
    """Row arrays x, v, a (n x 3) and t (n), NaN-filled past the cursor ``n`` (molecule.py:111-167)."""

    def __init__(self, beamline=None, n_rows: int | None = None):
        if n_rows is None:
            n_rows = 10 + sum(el.N_steps() for el in beamline.elements)
        self.x = np.full((n_rows, 3), np.nan)
        self.v = np.full((n_rows, 3), np.nan)
        self.a = np.full((n_rows, 3), np.nan)
        self.t = np.full((n_rows,), np.nan)
        self.n = 0

    @classmethod
    def from_rows(cls, rows: np.ndarray) -> "Trajectory":
        """Build from recorded rows [n][10] = x, y, z, vx, vy, vz, ax, ay, az, t."""
        tr = cls(n_rows=0)
        tr.x = np.ascontiguousarray(rows[:, 0:3])
        tr.v = np.ascontiguousarray(rows[:, 3:6])
        tr.a = np.ascontiguousarray(rows[:, 6:9])
        tr.t = np.ascontiguousarray(rows[:, 9])
        tr.n = rows.shape[0]
        return tr

    def update(self, x, v, a, t):
        if self.n >= self.x.shape[0]:
            self._grow(max(8, self.x.shape[0]))
        self.x[self.n, :] = x
        self.v[self.n, :] = v
        self.a[self.n, :] = a
        self.t[self.n] = t
        self.n += 1

    def _grow(self, k: int):
        self.x = np.concatenate((self.x, np.full((k, 3), np.nan)))
        self.v = np.concatenate((self.v, np.full((k, 3), np.nan)))
        self.a = np.concatenate((self.a, np.full((k, 3), np.nan)))
        self.t = np.concatenate((self.t, np.full((k,), np.nan)))

    def add_steps(self, beamline):
        self._grow(sum(el.N_steps() for el in beamline.elements))

    def drop_nans(self):
        self.x = self.x[np.all(np.isfinite(self.x), axis=1), :]
        self.v = self.v[np.all(np.isfinite(self.v), axis=1), :]
        self.a = self.a[np.all(np.isfinite(self.a), axis=1), :]
        self.t = self.t[np.isfinite(self.t)]

    def save_to_hdf(self, file, run_name: str, group_name: str) -> None:
        self.drop_nans()
        group_path = run_name + "/" + group_name
        file.create_group(group_path)
        file[group_path].create_dataset("x", data=self.x)
        file[group_path].create_dataset("v", data=self.v)
        file[group_path].create_dataset("a", data=self.a)
        file[group_path].create_dataset("t", data=self.t)


def propagate_molecules_through(elements, molecules, g=None, set_detected=True, engine=None):
    """Run ``molecules`` through ``elements`` on the device, starting from each molecule's current
    row (position, velocity, stored acceleration, time), and append the rows the reference would
    append.  A stopped molecule is set dead with the stopping label; a survivor is labelled
    "Detected" only if ``set_detected`` (a whole-beamline walk, beamline.py:33-35)."""
    from .engine import Engine
    from .compile import compile_elements

    if not molecules:
        return
    g = float(_G if g is None else g)
    x = np.array([m.x() for m in molecules], dtype=np.float64).reshape(-1, 3)
    v = np.array([m.v() for m in molecules], dtype=np.float64).reshape(-1, 3)
    a = np.array([m.a() for m in molecules], dtype=np.float64).reshape(-1, 3)
    t = np.array([m.t() for m in molecules], dtype=np.float64)
    if engine is None:
        engine = Engine(compile_elements(list(elements), g))
    vz = v[:, 2]
    vz_min = float(np.min(vz[vz > 0])) if np.any(vz > 0) else 1.0
    cap = 1 + sum(el.rows_needed(vz_min) if hasattr(el, "rows_needed") else el.N_steps() for el in elements)
    rows, nrows, labels = engine.record_rows(x[:, 0], x[:, 1], x[:, 2], v[:, 0], v[:, 1], v[:, 2], t,
                                             a0=(a[:, 0], a[:, 1], a[:, 2]), cap=cap)
    names = engine.labels
    det = engine.cb.label_detected
    for i, m in enumerate(molecules):
        for r in rows[i, 1:nrows[i]]:
            m.trajectory.update(r[0:3], r[3:6], r[6:9], r[9])
        lab = int(labels[i])
        if lab == det:
            if set_detected:
                m.set_aperture_hit(names[lab])
        else:
            m.set_dead()
            m.set_aperture_hit(names[lab])
