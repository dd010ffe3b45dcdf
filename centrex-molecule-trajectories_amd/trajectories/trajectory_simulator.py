"""TrajectorySimulator, Counter and SimulationResult (trajectory_simulator.py:22-254 of the reference).

``run_simulation`` keeps the reference's signature and side effects.  The per-molecule joblib loop
(:52-83) is replaced by the device walker: the inputs are drawn on the host exactly as the
reference draws them (``N_loops = 100*n_jobs`` chunks of ``N = int(N_traj/N_loops)`` molecules, each
chunk drawing vdist then xdist), propagated on the GPU in HBM-resident chunks, and tallied into a
Counter whose key order is the reference's first-seen order.  Kept molecules
(``apertures_of_interest``) are re-simulated on demand by the record kernel to materialise their
Molecule objects and trajectories.
"""
from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass
from pathlib import Path
from typing import List

import warnings

import numpy as np

from . import _native as N
from .beamline import Beamline
from .distributions import CeNTREXPositionDistribution, CeNTREXVelocityDistribution, Distribution
from .engine import Engine
from .molecule import Molecule, Trajectory


def reference_chunking(N_traj: int, n_jobs: int):
    """(N_loops, N, total) of run_simulation: N_loops = 100*n_jobs chunks of N = int(N_traj/N_loops)
    molecules, so total = N*N_loops are simulated (trajectory_simulator.py:48-49)."""
    N_loops = 100 * n_jobs
    N = int(N_traj / N_loops)
    return N_loops, N, N * N_loops


class Counter:
    """Label -> count, insertion-ordered (trajectory_simulator.py:106-158)."""

    def __init__(self) -> None:
        self.counter_dict = {}

    def increment_counter(self, aperture_hit: str) -> None:
        self.counter_dict[aperture_hit] = self.counter_dict.get(aperture_hit, 0) + 1

    def print(self) -> None:
        print("Number of molecules that hit each element:")
        for key, value in self.counter_dict.items():
            print(f"{key} : {value}")

    def calculate_efficiency(self) -> float:
        total = sum(self.counter_dict.values())
        if "Detected" in self.counter_dict:
            return self.counter_dict["Detected"] / total
        return 0

    def merge_counters(self, others: List) -> None:
        for other in others:
            for key, value in other.counter_dict.items():
                self.counter_dict[key] = self.counter_dict.get(key, 0) + value

    def save_to_hdf(self, filepath: Path, run_name: str) -> None:
        from .hdf import save_counter

        save_counter(self, filepath, run_name)

    @classmethod
    def from_tallies(cls, labels, counts, first_index) -> "Counter":
        """Counter with the labels that occurred, in order of the first molecule reaching each."""
        c = cls()
        order = sorted((int(first_index[i]), i) for i in range(len(labels)) if counts[i] > 0)
        for _, i in order:
            c.counter_dict[labels[i]] = int(counts[i])
        return c


class MoleculeList(Sequence):
    """Kept molecules, materialised on access.

    Holds the kept molecules' initial conditions and labels; indexing re-runs them (in blocks) with
    the record kernel -- the propagation is deterministic, so the trajectories are the ones the
    simulation produced."""

    BLOCK = 2048

    def __init__(self, engine: Engine, init: np.ndarray, labels: np.ndarray, cap: int, states=None):
        self._engine = engine
        self._init = init  # [7][n]: x, y, z, vx, vy, vz, t
        self._labels = labels
        self._states = None if states is None else np.ascontiguousarray(states, dtype=np.uint8)
        self._cap = cap
        self._cache = {}

    def __len__(self):
        return self._init.shape[1]

    @property
    def aperture_hits(self) -> list:
        """The stopping label of every kept molecule, in order (without materialising them)."""
        names = self._engine.labels
        return [names[int(k)] for k in self._labels]

    def sample_at(self, z: float, select=None):
        """State of the kept molecules ``select`` (all if None) at ``z`` by the post-processing rule
        (post_processing.py:20-140): one device launch that re-walks them with the z-grid recorder.
        Returns ([6][m] x, y, vx, vy, vz, t; NaN where a molecule ends before z, valid [m])."""
        idx = np.arange(len(self)) if select is None else np.asarray(select, dtype=np.int64)
        if idx.size == 0:
            return np.zeros((6, 0)), np.zeros(0, bool)
        I = self._init[:, idx]
        st = None if self._states is None else self._states[idx]
        data, written, lab = self._engine.zgrid(*I, state=st, z_start=float(z), z_step=1.0, n_samples=1)
        if not np.array_equal(lab, self._labels[idx]):
            raise RuntimeError("re-walked kept molecules stopped at other labels than in the run")
        return np.ascontiguousarray(data[:, 0, :]), written.astype(bool)

    def _block(self, b):
        if b not in self._cache:
            if len(self._cache) > 8:
                self._cache.clear()
            s = slice(b * self.BLOCK, min((b + 1) * self.BLOCK, len(self)))
            I = self._init[:, s]
            st = None if self._states is None else self._states[s]
            rows, nrows, lab = self._engine.record_rows(*I, state=st, cap=self._cap)
            if not np.array_equal(lab, self._labels[s]):
                raise RuntimeError("re-recorded kept molecules stopped at other labels than in the run")
            self._cap = max(self._cap, int(nrows.max()) if nrows.size else 0)
            names = self._engine.labels
            mols = []
            for j in range(rows.shape[0]):
                m = Molecule()
                m.g = self._engine.cb.phys.g
                m.trajectory = Trajectory.from_rows(rows[j, : nrows[j]])
                m.trajectory.drop_nans()
                m.set_aperture_hit(names[int(lab[j])])
                if int(lab[j]) != self._engine.cb.label_detected:
                    m.set_dead()
                mols.append(m)
            self._cache[b] = mols
        return self._cache[b]

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[k] for k in range(*i.indices(len(self)))]
        if i < 0:
            i += len(self)
        if not 0 <= i < len(self):
            raise IndexError(i)
        return self._block(i // self.BLOCK)[i % self.BLOCK]


class TrajectorySimulator:
    def __init__(self) -> None:
        self.counter = Counter()
        self.results = {}

    def run_simulation(self, beamline: Beamline, run_name: str, vdist=CeNTREXVelocityDistribution(),
                       xdist=CeNTREXPositionDistribution(), N_traj: int = 1000, apertures_of_interest=[],
                       n_jobs=1, *, ctx=None, chunk: int = 1 << 22, block: int = 0, states=None,
                       raise_out_of_table: bool = True, device_seed: int | None = None):
        """Simulate ``int(N_traj/(100*n_jobs))*100*n_jobs`` molecules (trajectory_simulator.py:34-103).

        ``n_jobs`` only sets the reference's chunking (and hence which molecules are drawn); the
        work runs on the GPU.  ``states`` (extension) gives each molecule a lens state id.
        ``device_seed`` (extension) draws the molecules on the device from counter-based streams
        (``simulate_device``) instead of replaying numpy's global RNG on the host: the same
        distributions, other samples, and no host draw or upload."""
        N_loops, N, total = reference_chunking(N_traj, n_jobs)
        if device_seed is not None:
            res = simulate_device(beamline, xdist, vdist, total, device_seed, apertures_of_interest, ctx=ctx,
                                  chunk=chunk, block=block, states=states, raise_out_of_table=raise_out_of_table)
            self._store(run_name, beamline, xdist, vdist, res)
            return
        vs, xs = [], []
        for _ in range(N_loops):
            vs.append(np.asarray(vdist.draw(N), dtype=np.float64).reshape(3, N))
            xs.append(np.asarray(xdist.draw(N), dtype=np.float64).reshape(3, N))
        V = np.concatenate(vs, axis=1) if vs else np.zeros((3, 0))
        X = np.concatenate(xs, axis=1) if xs else np.zeros((3, 0))
        res = simulate_arrays(beamline, X, V, apertures_of_interest, ctx=ctx, chunk=chunk, block=block,
                              states=states, raise_out_of_table=raise_out_of_table)
        assert res.n == total
        self._store(run_name, beamline, xdist, vdist, res)

    def _store(self, run_name, beamline, xdist, vdist, res):
        self.counter = res.counter
        self.result = SimulationResult(self.counter, beamline, xdist, vdist, res.molecules)
        self.result.run = res
        self.results[run_name] = SimulationResult(self.counter, beamline, xdist, vdist, res.molecules)
        self.results[run_name].run = res


@dataclass
class RunOutput:
    n: int
    labels: list
    tally: np.ndarray  # [n_states][n_labels]
    first_index: np.ndarray
    counter: Counter
    aoi_index: np.ndarray
    aoi_label: np.ndarray
    aoi_final: dict  # final phase space of kept molecules
    molecules: MoleculeList
    work: np.ndarray
    n_out_of_table: int
    n_capped: int = 0


def simulate_arrays(beamline: Beamline, X: np.ndarray, V: np.ndarray, apertures_of_interest=(), *, ctx=None,
                    chunk: int = 1 << 22, block: int = 0, states=None, t0=None, raise_out_of_table=True,
                    engine: Engine | None = None) -> RunOutput:
    """Propagate molecules with initial positions X (3, n) and velocities V (3, n) on the device."""
    eng = engine or Engine(beamline, ctx, n_states=None if states is None else int(np.max(states)) + 1)
    n = X.shape[1]
    nl, ns = eng.cb.n_labels, eng.cb.n_states
    tally = np.zeros((ns, nl), np.int64)
    first = np.full(nl, np.iinfo(np.int64).max, np.int64)
    work = np.zeros(8, np.uint64)
    aoi_idx, aoi_lab, aoi_fin = [], [], []
    n_oot = n_capped = 0
    for s in range(0, max(n, 1), chunk):
        e = min(n, s + chunk)
        if e <= s:
            break
        r = eng.propagate(X[0, s:e], X[1, s:e], X[2, s:e], V[0, s:e], V[1, s:e], V[2, s:e],
                          None if t0 is None else t0[s:e], state=None if states is None else states[s:e],
                          outputs=False, aoi_labels=apertures_of_interest, index_base=s, block=block,
                          per_molecule=False)
        tally += r["tally"]
        first = np.minimum(first, r["first_index"])
        work += r["work"]
        n_oot += int(r["work"][N.WORK_OOT])
        n_capped += int(r["work"][N.WORK_CAPPED])
        if "aoi_index" in r:
            o = np.argsort(r["aoi_index"], kind="stable")
            aoi_idx.append(r["aoi_index"][o])
            aoi_lab.append(r["aoi_label"][o])
            aoi_fin.append({k: v[o] for k, v in r["aoi_out"].items()})
    eng.free_scratch()
    idx = np.concatenate(aoi_idx) if aoi_idx else np.zeros(0, np.int64)
    init = np.zeros((7, idx.size))
    if idx.size:
        init[0:3] = X[:, idx]
        init[3:6] = V[:, idx]
        init[6] = 0.0 if t0 is None else t0[idx]
    return _run_output(eng, beamline, n, tally, first, work, n_oot, n_capped, aoi_idx, aoi_lab, aoi_fin, init,
                       states, raise_out_of_table)


def _run_output(eng, beamline, n, tally, first, work, n_oot, n_capped, aoi_idx, aoi_lab, aoi_fin, init, states,
                raise_out_of_table) -> RunOutput:
    """The RunOutput of a chunked run: the out-of-table error or step-cap warning, the Counter, and the
    kept molecules (global indices aoi_idx, initial conditions init [7][k]) as a MoleculeList."""
    if n_oot and raise_out_of_table:
        raise ValueError(f"{n_oot} molecules evaluated the lens force outside its table "
                         "(the reference's interp1d raises ValueError here)")
    if n_capped:
        warnings.warn(f"{n_capped} molecules needed more than max_lens_steps full RK4 steps in a lens; their "
                      "step count was cut there and one partial step covered the rest of the lens", RuntimeWarning)
    idx = np.concatenate(aoi_idx) if aoi_idx else np.zeros(0, np.int64)
    lab = np.concatenate(aoi_lab) if aoi_lab else np.zeros(0, np.int32)
    fin = {k: np.concatenate([d[k] for d in aoi_fin]) for k in aoi_fin[0]} if aoi_fin else {}
    counter = Counter.from_tallies(eng.cb.labels, tally.sum(axis=0), first)
    vz = init[5] if idx.size else np.array([1.0])
    vzmin = float(vz[vz > 0].min()) if np.any(vz > 0) else 1.0
    cap = 10 + sum(el.rows_needed(vzmin) if hasattr(el, "rows_needed") else el.N_steps() for el in beamline.elements)
    kept_states = None if states is None else np.asarray(states, dtype=np.uint8)[idx]
    mols = MoleculeList(eng, init, lab, cap, states=kept_states)
    return RunOutput(n, list(eng.cb.labels), tally, first, counter, idx, lab, fin, mols, work, n_oot, n_capped)


def device_stream_keys(seed: int) -> tuple[int, int]:
    """Keys of the position and velocity streams of a device-drawn run: 2 seed and 2 seed + 1 (mod 2^64),
    so that the two never share a stream even where both are Gaussian."""
    k = (2 * int(seed)) % (1 << 64)
    return k, (k + 1) % (1 << 64)


def simulate_device(beamline: Beamline, xdist: Distribution, vdist: Distribution, n: int, seed: int,
                    apertures_of_interest=(), *, ctx=None, chunk: int = 1 << 22, block: int = 0, states=None,
                    raise_out_of_table=True, engine: Engine | None = None) -> RunOutput:
    """Propagate n molecules drawn on the device: positions by ``xdist.draw_device`` and velocities by
    ``vdist.draw_device`` under the keys of ``device_stream_keys(seed)``, molecule i taking counter i in
    both, so the molecules do not depend on ``chunk``.  The inputs never leave the device; the kept
    molecules' initial conditions are gathered there.  Results are those of ``simulate_arrays`` on the
    same draws copied to the host."""
    from .engine import FIELDS, gather

    eng = engine or Engine(beamline, ctx, n_states=None if states is None else int(np.max(states)) + 1)
    ctx = eng.ctx
    kx, kv = device_stream_keys(seed)
    nl, ns = eng.cb.n_labels, eng.cb.n_states
    tally = np.zeros((ns, nl), np.int64)
    first = np.full(nl, np.iinfo(np.int64).max, np.int64)
    work = np.zeros(8, np.uint64)
    aoi_idx, aoi_lab, aoi_fin, inits = [], [], [], []
    n_oot = n_capped = 0
    bufs = [ctx.alloc(max(1, min(chunk, n))) for _ in range(6)]
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        src = [*xdist.draw_device(m, kx, offset=s, ctx=ctx, out=bufs[:3]),
               *vdist.draw_device(m, kv, offset=s, ctx=ctx, out=bufs[3:])]
        st = None
        if states is not None:
            sh = np.ascontiguousarray(states[s:s + m], dtype=np.uint8)
            if m and int(sh.max()) >= ns:
                raise ValueError("state id >= n_states")
            st = ctx.to_device(sh)
        r = eng.propagate_device(m, dict(zip(FIELDS[:6], src)), None, st, outputs=False,
                                 aoi_labels=apertures_of_interest, index_base=s, block=block, per_molecule=False)
        tally += r["tally"]
        first = np.minimum(first, r["first_index"])
        work += r["work"]
        n_oot += int(r["work"][N.WORK_OOT])
        n_capped += int(r["work"][N.WORK_CAPPED])
        if "aoi_index" in r:
            o = np.argsort(r["aoi_index"], kind="stable")
            idx = np.ascontiguousarray(r["aoi_index"][o], dtype=np.int64)
            aoi_idx.append(idx)
            aoi_lab.append(r["aoi_label"][o])
            aoi_fin.append({k: v[o] for k, v in r["aoi_out"].items()})
            if idx.size:
                dst = [ctx.alloc(idx.size) for _ in range(6)]
                gather(ctx, idx.size, ctx.to_device(idx), s, m, src, dst)
                inits.append(np.array([d.to_host()[: idx.size] for d in dst]))
    eng.free_scratch()
    for b in bufs:
        b.free()
    k = sum(a.size for a in aoi_idx)
    init = np.zeros((7, k))
    if k:
        init[0:6] = np.concatenate(inits, axis=1)
    return _run_output(eng, beamline, n, tally, first, work, n_oot, n_capped, aoi_idx, aoi_lab, aoi_fin, init,
                       states, raise_out_of_table)


@dataclass
class SimulationResult:
    counter: Counter
    beamline: Beamline
    xdist: Distribution
    vdist: Distribution
    molecules: List[Molecule]

    def plot(self, N_max: int = 10000, elements: List[str] = None, show: bool = True):
        import matplotlib.pyplot as plt

        axes = self.beamline.plot()
        n_max = min(N_max, len(self.molecules))
        n = i = 0
        while n < n_max and i < len(self.molecules):
            molecule = self.molecules[i]
            i += 1
            if elements is not None and molecule.aperture_hit not in elements:
                continue
            molecule.plot_trajectory(axes)
            n += 1
        if show:
            plt.show()
        return axes

    def save_to_hdf(self, filepath: Path, run_name: str) -> None:
        from .hdf import save_result

        save_result(self, filepath, run_name)
