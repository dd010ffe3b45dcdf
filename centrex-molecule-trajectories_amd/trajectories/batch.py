"""Batched scan: many beamline configurations in one persistent, graph-replayed launch.

The reference runs ``TrajectorySimulator.run_simulation`` (trajectory_simulator.py:34-103) once per
configuration.  ``BatchScan`` packs all configurations (their own element lists, molecule counts and
Gaussian sources) into one ``ctraj_batch`` and runs every molecule of every configuration in one
persistent kernel; ``autotune`` picks the block size, the queue chunk and the lane-refill threshold
by timing candidates on the device.  Molecule j of configuration c is molecule j of
``ctraj_draw_gaussian(seed_c)`` over (x, y, z, vx, vy, vz), so each configuration's tallies equal a
``ctraj_propagate`` run over those draws.
"""
from __future__ import annotations

import ctypes as C
import itertools
import json
from pathlib import Path

import numpy as np

from . import _native as N
from .compile import LABEL_DETECTED, Registry
from .device import Context, get_context


class BatchScan:
    """``beamlines``: list of Beamline (one per configuration); ``counts``: molecules per
    configuration; ``seeds``: device source seed per configuration; ``mean``/``sigma``: the Gaussian
    source (x, y, z, vx, vy, vz), shared or one row per configuration."""

    def __init__(self, beamlines, counts, seeds, mean, sigma, *, ctx: Context | None = None,
                 max_lens_steps: int = 0):
        self.ctx = ctx or get_context()
        self.beamlines = list(beamlines)
        n_cfg = len(self.beamlines)
        counts = np.asarray(counts, dtype=np.int64)
        seeds = np.asarray(seeds, dtype=np.uint64)
        mean = np.broadcast_to(np.asarray(mean, dtype=np.float64), (n_cfg, 6))
        sigma = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (n_cfg, 6))
        if counts.shape != (n_cfg,) or seeds.shape != (n_cfg,):
            raise ValueError("one count and one seed per configuration")
        g = {bl.g for bl in self.beamlines}
        if len(g) != 1:
            raise ValueError("all configurations must share g")
        reg = Registry()
        packed, slices = [], []
        for bl in self.beamlines:
            start = len(packed)
            for el in bl.elements:
                packed.append(el._ctraj_pack(reg))
            slices.append((start, len(packed) - start))
        det = reg.label(LABEL_DETECTED)
        self.labels = list(reg.labels)
        self.label_detected = det
        self.counts = counts
        self.seeds = seeds
        arr = (N.Elem * max(1, len(packed)))(*packed)
        cfgs = (N.BatchCfg * n_cfg)()
        for k, (s, ne) in enumerate(slices):
            cfgs[k].elem_begin, cfgs[k].n_elems = s, ne
            cfgs[k].n = int(counts[k])
            cfgs[k].seed = int(seeds[k])
            cfgs[k].mean[:] = [float(v) for v in mean[k]]
            cfgs[k].sigma[:] = [float(v) for v in sigma[k]]
        self.phys = N.Phys(g=float(g.pop()), n_labels=len(self.labels), label_detected=det, n_states=1,
                           max_lens_steps=int(max_lens_steps))
        L = N.lib()
        if reg.wsplines:
            coef = np.ascontiguousarray(np.stack([s.coef for s in reg.wsplines[:1]]), dtype=np.float64)
            N.check("ctraj_set_wspline", L.ctraj_set_wspline(self.ctx.handle, 1, coef.shape[1], coef.ctypes.data),
                    self.ctx.handle)
            self.ctx.bound = None  # the context's tables changed: propagation must re-bind its beamline
        h = C.c_void_p()
        N.check("ctraj_batch_create", L.ctraj_batch_create(self.ctx.handle, arr, len(packed), cfgs, n_cfg,
                                                           C.byref(self.phys), C.byref(h)), self.ctx.handle)
        self.handle = h.value
        self._wspline = reg.wsplines[0] if reg.wsplines else None
        self.n_cfg = n_cfg
        self.total = int(counts.sum())
        self.tune = N.BatchTune(block=0, chunk=128, grid=0, refill=0, use_graph=1, fused=0)

    @property
    def split_path(self) -> bool:
        """True when fused = 0 runs the split kernels (walk and queue, lens, walk after the lens)."""
        return bool(N.lib().ctraj_batch_split_path(self.handle))

    def run(self, tune: N.BatchTune | None = None, graph: bool | None = None):
        """One launch over every configuration.  Returns (tally [n_cfg][n_labels], work, device ms)."""
        t = N.BatchTune()
        C.memmove(C.byref(t), C.byref(tune or self.tune), C.sizeof(t))
        if graph is not None:
            t.use_graph = int(graph)
        tally = np.zeros((self.n_cfg, len(self.labels)), np.int64)
        work = np.zeros(N.WORK_N, np.uint64)
        ms = C.c_float()
        N.check("ctraj_batch_run", N.lib().ctraj_batch_run(self.handle, C.byref(t), tally.ctypes.data,
                                                           work.ctypes.data, C.byref(ms)), self.ctx.handle)
        return tally, work, ms.value

    def detected_fraction(self, tally: np.ndarray) -> np.ndarray:
        return tally[:, self.label_detected] / np.maximum(self.counts, 1)

    def signature(self) -> str:
        return f"cfg{self.n_cfg}-n{self.total}-labels{len(self.labels)}"

    def autotune(self, blocks=(256, 512, 1024), chunks=(128,), refills=(8, 12, 16, 24), fused=(0,), reps: int = 2,
                 cache: str | Path | None = None, verbose: bool = False) -> N.BatchTune:
        """Time every (block, chunk, refill, fused) candidate (median of ``reps`` device-timed launches
        after one warm-up) and keep the fastest as ``self.tune`` (graph replay on).  ``fused`` lists the
        paths to try: 0 the split kernels (``block`` is the lens kernel's workgroup, up to 1024; ``chunk``
        is unused), 1 the single-kernel walker (``block`` up to 256).  With ``cache`` (a JSON file), a
        previous result for the same device and batch signature is reused."""
        key = f"{self.ctx.arch}:{self.signature()}"
        if cache is not None:
            p = Path(cache)
            if p.exists():
                known = json.loads(p.read_text())
                if key in known:
                    k = known[key]
                    self.tune = N.BatchTune(block=k["block"], chunk=k["chunk"], grid=0, refill=k["refill"],
                                            use_graph=1, fused=k.get("fused", 0))
                    return self.tune
        results = []
        for f, b, c, r in itertools.product(fused, blocks, chunks, refills):
            if (f and b > 256) or (not f and c != chunks[0]):  # the split kernels take no chunk
                continue
            t = N.BatchTune(block=b, chunk=c, grid=0, refill=r, use_graph=0, fused=f)
            self.run(t)
            times = sorted(self.run(t)[2] for _ in range(reps))
            results.append((times[len(times) // 2], b, c, r, f))
            if verbose:
                print(f"fused {f} block {b:4d} chunk {c:4d} refill {r:2d}: {times[len(times) // 2]:.3f} ms", flush=True)
        results.sort()
        _, b, c, r, f = results[0]
        self.tune = N.BatchTune(block=b, chunk=c, grid=0, refill=r, use_graph=1, fused=f)
        self.tuning = results
        if cache is not None:
            p = Path(cache)
            known = json.loads(p.read_text()) if p.exists() else {}
            known[key] = {"block": b, "chunk": c, "refill": r, "fused": f, "ms": results[0][0]}
            p.write_text(json.dumps(known, indent=1))
        return self.tune

    def close(self):
        if getattr(self, "handle", None):
            N.lib().ctraj_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
