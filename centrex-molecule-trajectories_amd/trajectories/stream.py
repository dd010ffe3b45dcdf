"""Streamed large runs: molecules drawn on the device chunk by chunk, per-state tallies, and the
z-grid trajectories of every detected molecule in one dense device buffer.

This is ``TrajectorySimulator.run_simulation`` (trajectory_simulator.py:34-103) for runs too large
to draw on the host, with the reference's saved trajectories (trajectory_simulator.py:75-76) resampled
every ``z_step`` by the post-processing rule (post_processing.py:20-78).

Pass 1, per chunk of ``chunk`` molecules kept in HBM: draw (x, y, z, vx, vy, vz) with the device
Gaussian source at offset = the chunk's first molecule and a uniform lens-state id per molecule;
propagate (per-state, per-element tallies and the global indices of detected molecules
accumulate on the device; the host never waits between chunks).
Pass 2: sort the detected indices (radix sort), regenerate those molecules' inputs from the same
counter-based streams, and re-walk them in one record launch that writes their samples
[n_detected][n_samples][6] (x, y, vx, vy, vz, t): each molecule's samples contiguous, so that every
lane writes one sequential stream (the lanes of a wavefront soon sample different z).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .device import Context, DeviceArray, get_context
from .engine import (Engine, _soa, bind, draw_gaussian, draw_gaussian_at, draw_uniform_u8, draw_uniform_u8_at,
                     gather, sort_i64)

FIELDS6 = ("x", "y", "z", "vx", "vy", "vz")


@dataclass
class StreamResult:
    n: int
    labels: list
    tally: np.ndarray               # [n_states][n_labels]
    detected_index: np.ndarray      # global molecule index of each detected molecule, ascending
    detected_state: np.ndarray
    n_samples: int
    z_start: float
    z_step: float
    trajectories: DeviceArray | None  # [n_detected][n_samples][6] on the device, or None
    samples_written: np.ndarray | None
    timings: dict = field(default_factory=dict)
    work: np.ndarray | None = None

    @property
    def n_detected(self) -> int:
        return int(self.detected_index.size)

    def trajectories_of(self, js) -> np.ndarray:
        """Samples of the detected molecules ``js`` (positions in ``detected_index``) as a host array
        [len(js)][n_samples][6] (x, y, vx, vy, vz, t), gathered on the device."""
        if self.trajectories is None:
            raise ValueError("trajectories were not recorded")
        ctx = self.trajectories.ctx
        js = np.atleast_1d(np.asarray(js, dtype=np.int64))
        nd, ns = self.n_detected, self.n_samples
        if js.size and (js.min() < 0 or js.max() >= nd):
            raise IndexError("detected molecule out of range")
        e = np.arange(ns * 6)[None, :]
        idx = np.ascontiguousarray((js[:, None] * (ns * 6) + e).reshape(-1))
        d_idx = ctx.to_device(idx)
        out = ctx.alloc(max(idx.size, 1))
        gather(ctx, idx.size, d_idx, 0, 6 * ns * nd, [self.trajectories], [out])
        return out.to_host()[: idx.size].reshape(js.size, ns, 6)


class StreamedRun:
    def __init__(self, beamline, n: int, *, chunk: int = 1 << 25, seed: int = 4, n_states: int = 4,
                 mean=None, sigma=None, z_start: float = 0.0, z_step: float = 1e-3, z_end: float | None = None,
                 ctx: Context | None = None):
        from . import configs

        self.ctx = ctx or get_context()
        self.engine = Engine(beamline, self.ctx, n_states=n_states)
        self.n = int(n)
        self.chunk = int(chunk)
        self.seed = int(seed)
        self.n_states = int(n_states)
        self.mean = np.asarray(configs.SOURCE_MEAN if mean is None else mean, dtype=np.float64)
        self.sigma = np.asarray(configs.SOURCE_SIGMA if sigma is None else sigma, dtype=np.float64)
        self.z_start = float(z_start)
        self.z_step = float(z_step)
        if z_end is None:
            z_end = max(el.z1 for el in beamline.elements)
        self.n_samples = int(np.floor((z_end - self.z_start) / self.z_step + 1e-9)) + 1

    def run(self, trajectories: bool = True) -> StreamResult:
        ctx, eng = self.ctx, self.engine
        cb = eng.cb
        nl, ns = cb.n_labels, cb.n_states
        det = cb.label_detected
        ch = min(self.chunk, max(self.n, 1))
        inp = {k: ctx.alloc(ch) for k in FIELDS6}
        state = ctx.alloc(ch, np.uint8)
        tally = ctx.alloc(ns * nl, np.int64)
        work = ctx.alloc(N.WORK_N, np.uint64)
        aoi_count = ctx.alloc(1, np.int64)
        cap = max(4096, int(0.05 * self.n))
        while True:
            aoi_index = ctx.alloc(cap, np.int64)
            for a in (tally, work, aoi_count):
                a.fill(0)
            t0 = time.perf_counter()
            ctx.event_record(0)
            # pass 1: no host synchronisation; kept indices accumulate on the device
            for start in range(0, self.n, ch):
                m = min(ch, self.n - start)
                draw_gaussian(ctx, m, self.mean, self.sigma, self.seed, offset=start, out=[inp[k] for k in FIELDS6])
                draw_uniform_u8(ctx, m, self.n_states, self.seed, offset=start, out=state)
                eng.launch(inp, state=state, tally=tally, work=work, aoi_mask=1 << det, aoi_cap=cap,
                           aoi_count=aoi_count, aoi_index=aoi_index, n=m, index_base=start)
            ctx.event_record(1)
            count = int(aoi_count.to_host()[0])
            wall1 = time.perf_counter() - t0
            t_pass1 = ctx.event_elapsed_ms(0, 1)
            if count <= cap:
                break
            cap = count  # the kept molecules outgrew the index store: run again with room for all
        traj = written = None
        t_pass2 = 0.0
        st_h = np.zeros(0, np.uint8)
        idx_h = np.zeros(0, np.int64)
        t_alloc = 0.0
        if count:
            # output buffers first, so that the timed device region holds pass 2's kernels only
            t0 = time.perf_counter()
            init = [ctx.alloc(count) for _ in FIELDS6]
            init_state = ctx.alloc(count, np.uint8)
            if trajectories:
                traj = ctx.alloc(6 * self.n_samples * count)
                written = ctx.alloc(count, np.int32)
            t_alloc = time.perf_counter() - t0
            ctx.event_record(2)
            sort_i64(ctx, aoi_index, count)
            draw_gaussian_at(ctx, aoi_index, count, self.mean, self.sigma, self.seed, out=init)
            draw_uniform_u8_at(ctx, aoi_index, count, self.n_states, self.seed, out=init_state)
            if trajectories:
                R = N.Record()
                R.n = count
                R.index = None
                R.inp = _soa(dict(zip(FIELDS6, init)))
                R.z_const, R.t_const = 0.0, 0.0
                R.state = init_state.ptr
                R.mode, R.cap = N.REC_ZGRID_MOL, self.n_samples
                R.z_start, R.z_step = self.z_start, self.z_step
                R.buf, R.n_rows, R.label = traj.ptr, written.ptr, None
                bind(ctx, cb)
                N.check("ctraj_record_trajectories", N.lib().ctraj_record_trajectories(ctx.handle, C.byref(R)),
                        ctx.handle)
            ctx.event_record(3)
            ctx.sync()
            t_pass2 = ctx.event_elapsed_ms(2, 3)
            idx_h = np.empty(count, np.int64)
            ctx.memcpy(idx_h.ctypes.data, aoi_index.ptr, idx_h.nbytes, N.D2H)
            st_h = np.empty(count, np.uint8)
            ctx.memcpy(st_h.ctypes.data, init_state.ptr, st_h.nbytes, N.D2H)
            ctx.sync()
        return StreamResult(self.n, list(cb.labels), tally.to_host().reshape(ns, nl), idx_h, st_h, self.n_samples,
                            self.z_start, self.z_step, traj, None if written is None else written.to_host(),
                            {"pass1_device_ms": t_pass1, "pass1_wall_s": wall1, "pass2_device_ms": t_pass2,
                             "pass2_alloc_s": t_alloc},
                            work.to_host())

    def _grow(self, arr: DeviceArray, used: int, new_n: int) -> DeviceArray:
        out = self.ctx.alloc(new_n, arr.dtype)
        if used:
            self.ctx.memcpy(out.ptr, arr.ptr, used * arr.dtype.itemsize, N.D2D)
        return out
