"""Positions and velocities of simulated molecules at a given z (post_processing.py:1-158 of the reference).

``find_radial_pos_dist`` / ``find_vel_dist`` keep the reference's signatures, its rule and its output
shapes ((m, 2) positions, (m, 3) velocities, molecules in result order).  The per-molecule Python loop
is replaced by one device launch over all molecules of the result:

* a result of ``TrajectorySimulator.run_simulation`` holds its kept molecules as initial conditions
  (``MoleculeList``); they are walked again by the z-grid recorder with one sample at z, so no
  trajectory row is materialised;
* any other list of ``Molecule`` objects (built by ``Beamline.propagate_through``, say) has its stored
  rows packed into one device buffer and evaluated by ``ctraj_sample_rows``.

Both apply the rule of post_processing.py:40-75: skip a molecule whose last row lies before z; return
the first row at exactly z; otherwise advance the row before the first row with z_row >= z to z by
``take_timestep`` with the row's stored acceleration.
"""
from __future__ import annotations

import ctypes as C
from typing import List

import numpy as np

from . import _native as N
from .device import get_context


def take_timestep(x0: np.ndarray, v0: np.ndarray, a0: np.ndarray, dt: float):
    """Position and velocity after dt at constant acceleration (post_processing.py:9-17)."""
    x = x0 + v0 * dt + a0 * dt ** 2 / 2
    v = v0 + a0 * dt
    return x, v


def _rows_of(molecules) -> tuple[np.ndarray, np.ndarray]:
    """Packed rows [total][10] (x, y, z, vx, vy, vz, ax, ay, az, t) and offsets [n+1] of Molecule objects,
    each trajectory cut at its cursor with its NaN rows dropped (molecule.py:160-167)."""
    parts, off = [], [0]
    for m in molecules:
        tr = m.trajectory
        k = min(int(tr.n), tr.x.shape[0])
        r = np.concatenate((tr.x[:k], tr.v[:k], tr.a[:k], np.asarray(tr.t[:k]).reshape(-1, 1)), axis=1)
        r = r[np.all(np.isfinite(r), axis=1)]
        parts.append(r)
        off.append(off[-1] + r.shape[0])
    rows = np.concatenate(parts) if parts else np.zeros((0, 10))
    return np.ascontiguousarray(rows, dtype=np.float64), np.asarray(off, dtype=np.int64)


def sample_rows(molecules, z: float, ctx=None) -> tuple[np.ndarray, np.ndarray]:
    """Evaluate stored trajectories at z on the device: ([6][n] x, y, vx, vy, vz, t; valid [n])."""
    ctx = ctx or get_context()
    rows, off = _rows_of(molecules)
    n = off.size - 1
    if n == 0:
        return np.zeros((6, 0)), np.zeros(0, bool)
    d_rows = ctx.to_device(rows.reshape(-1) if rows.size else np.zeros(10))
    d_off = ctx.to_device(off)
    outs = [ctx.alloc(n) for _ in range(6)]
    valid = ctx.alloc(n, np.uint8)
    S = N.Sample()
    S.n, S.row_off, S.n_rows, S.rows, S.z = n, d_off.ptr, rows.shape[0], d_rows.ptr, float(z)
    for f in range(6):
        S.out[f] = outs[f].ptr
    S.valid = valid.ptr
    N.check("ctraj_sample_rows", N.lib().ctraj_sample_rows(ctx.handle, C.byref(S)), ctx.handle)
    ctx.sync()
    return np.array([o.to_host() for o in outs]), valid.to_host().astype(bool)


def sample_at(result, z: float, elements: List[str] = None) -> tuple[np.ndarray, np.ndarray]:
    """The molecules of ``result`` at z: ([6][m] x, y, vx, vy, vz, t of the molecules the reference keeps,
    in result order; their indices into result.molecules)."""
    mols = result.molecules
    hits = mols.aperture_hits if hasattr(mols, "aperture_hits") else [m.aperture_hit for m in mols]
    sel = np.arange(len(hits), dtype=np.int64)
    if elements is not None:
        sel = np.array([i for i, h in enumerate(hits) if h in elements], dtype=np.int64)
    if hasattr(mols, "sample_at"):
        vals, valid = mols.sample_at(z, sel)
    else:
        vals, valid = sample_rows([mols[int(i)] for i in sel], z)
    return vals[:, valid], sel[valid]


def find_radial_pos_dist(result, z: float, elements: List[str] = None) -> np.ndarray:
    """x and y of the molecules in ``result`` at z, shape (m, 2) (post_processing.py:20-78)."""
    print("Determining radial positions...")
    vals, _ = sample_at(result, z, elements)
    return np.array(list(vals[0:2].T))


def find_vel_dist(result, z: float, elements: List[str] = None) -> np.ndarray:
    """vx, vy, vz of the molecules in ``result`` at z, shape (m, 3) (post_processing.py:81-140)."""
    print("Determining radial velocities...")
    vals, _ = sample_at(result, z, elements)
    return np.array(list(vals[2:5].T))


def plot_radial_pos_dist(result, z: float):
    """Placeholder kept from the reference (post_processing.py:143-149), which does nothing."""
    pass


def plot_radial_vel_dist(result, z: float):
    """Placeholder kept from the reference (post_processing.py:152-158), which does nothing."""
    pass
