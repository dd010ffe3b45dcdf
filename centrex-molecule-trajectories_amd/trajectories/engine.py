"""Host runner: binds a compiled beamline to a device context and launches the kernels.

``Engine.propagate`` / ``record_rows`` / ``zgrid`` take host numpy arrays (copied to and from the
device); ``Engine.launch`` takes device buffers and queues one kernel without synchronising -- the
form the benchmark times.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .compile import CompiledBeamline, compile_elements
from .device import Context, DeviceArray, get_context

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "t")


def bind(ctx: Context, cb: CompiledBeamline) -> None:
    """Make ``cb`` the beamline of ``ctx`` (uploads tables and the element table if changed)."""
    if ctx.bound is not None and ctx.bound == cb.key:
        return
    L = N.lib()
    for i, (r, a) in enumerate(cb.rtables):
        N.check("ctraj_set_rtable",
                L.ctraj_set_rtable(ctx.handle, i, r.ctypes.data, a.ctypes.data, int(r.size)), ctx.handle)
    coef = cb.wspline_coef()
    if coef is not None:
        N.check("ctraj_set_wspline",
                L.ctraj_set_wspline(ctx.handle, coef.shape[0], coef.shape[1], coef.ctypes.data), ctx.handle)
    N.check("ctraj_set_beamline",
            L.ctraj_set_beamline(ctx.handle, cb.elems, cb.n_elems, C.byref(cb.phys)), ctx.handle)
    ctx.bound = cb.key


def _soa(arrs) -> N.SoA:
    s = N.SoA()
    for k in FIELDS:
        v = arrs.get(k) if arrs else None
        setattr(s, k, N.addr(v))
    return s


class Engine:
    """A compiled beamline on a device context."""

    def __init__(self, beamline, ctx: Context | None = None, n_states: int | None = None, max_lens_steps: int = 0):
        self.ctx = ctx or get_context()
        if isinstance(beamline, CompiledBeamline):
            self.cb = beamline
        else:
            self.cb = compile_elements(beamline.elements, beamline.g, n_states=n_states,
                                       max_lens_steps=max_lens_steps)
        bind(self.ctx, self.cb)
        self._scratch = {}

    def _buf(self, name: str, n: int, dtype=np.float64) -> DeviceArray:
        """A device buffer of at least n elements kept for the host-array calls of this engine (each
        call copies its results to the host before it returns, so the next call may reuse it)."""
        b = self._scratch.get(name)
        if b is None or b.n < n or b.dtype != np.dtype(dtype):
            b = self.ctx.alloc(max(int(n), 1), dtype)
            self._scratch[name] = b
        return b

    def free_scratch(self) -> None:
        """Release the buffers kept by _buf (a chunked run calls this when it is done)."""
        for b in self._scratch.values():
            b.free()
        self._scratch.clear()

    @property
    def labels(self):
        return self.cb.labels

    # ------------------------------------------------------------------ device-buffer form
    def launch(self, inp: dict, *, out: dict | None = None, out_a=None, label=None, status=None, tally=None,
               first_index=None, aoi_mask: int = 0, aoi_cap: int = 0, aoi_count=None, aoi_index=None,
               aoi_label=None, aoi_out: dict | None = None, work=None, a0=None, state=None, n: int | None = None,
               index_base: int = 0, z_const: float = 0.0, t_const: float = 0.0, block: int = 0, grid: int = 0,
               flags: int = 0, refill: int = 0, noalloc: bool = False) -> None:
        """Queue one ``ctraj_propagate`` on device buffers (no synchronisation); ``noalloc`` calls
        ``ctraj_propagate_noalloc``, which allocates no device memory."""
        bind(self.ctx, self.cb)
        R = N.Run()
        R.n = int(n if n is not None else len(inp["x"]))
        R.index_base = int(index_base)
        R.inp = _soa(inp)
        R.z_const, R.t_const = float(z_const), float(t_const)
        for k in range(3):
            R.a0[k] = N.addr(a0[k]) if a0 is not None else None
            R.out_a[k] = N.addr(out_a[k]) if out_a is not None else None
        R.state = N.addr(state)
        R.out = _soa(out)
        R.label, R.status, R.tally, R.first_index = (N.addr(label), N.addr(status), N.addr(tally),
                                                    N.addr(first_index))
        R.aoi_mask, R.aoi_cap = int(aoi_mask), int(aoi_cap)
        R.aoi_count, R.aoi_index, R.aoi_label = N.addr(aoi_count), N.addr(aoi_index), N.addr(aoi_label)
        R.aoi_out = _soa(aoi_out)
        R.work = N.addr(work)
        R.block, R.grid, R.flags, R.refill = int(block), int(grid), int(flags), int(refill)
        if noalloc:
            N.check("ctraj_propagate_noalloc", N.lib().ctraj_propagate_noalloc(self.ctx.handle, C.byref(R)),
                    self.ctx.handle)
        else:
            N.check("ctraj_propagate", N.lib().ctraj_propagate(self.ctx.handle, C.byref(R)), self.ctx.handle)

    # ------------------------------------------------------------------ host-array form
    def _upload_inputs(self, x, y, z, vx, vy, vz, t, a0, state):
        ctx = self.ctx
        n = int(np.asarray(x).size)
        inp = {k: ctx.to_device(np.asarray(v, dtype=np.float64)) for k, v in
               zip(FIELDS[:6], (x, y, z, vx, vy, vz))}
        if t is not None:
            inp["t"] = ctx.to_device(np.asarray(t, dtype=np.float64))
        a0d = None
        if a0 is not None:
            a0d = [ctx.to_device(np.asarray(a, dtype=np.float64)) for a in a0]
        st = None
        if state is not None:
            st = ctx.to_device(np.asarray(state, dtype=np.uint8))
            if n and int(np.max(state)) >= self.cb.n_states:
                raise ValueError("state id >= n_states")
        return n, inp, a0d, st

    def propagate(self, x, y, z, vx, vy, vz, t=None, a0=None, state=None, *, outputs: bool = True,
                  want_accel: bool = False, aoi_labels=(), index_base: int = 0, block: int = 0, flags: int = 0,
                  refill: int = 0, noalloc: bool = False, per_molecule: bool = True) -> dict:
        """Propagate host arrays; returns labels, tallies, work counters and (optionally) final states."""
        n, inp, a0d, st = self._upload_inputs(x, y, z, vx, vy, vz, t, a0, state)
        return self.propagate_device(n, inp, a0d, st, outputs=outputs, want_accel=want_accel, aoi_labels=aoi_labels,
                                     index_base=index_base, block=block, flags=flags, refill=refill, noalloc=noalloc,
                                     per_molecule=per_molecule)

    def propagate_device(self, n: int, inp: dict, a0=None, state=None, *, outputs: bool = True,
                         want_accel: bool = False, aoi_labels=(), index_base: int = 0, block: int = 0, flags: int = 0,
                         refill: int = 0, noalloc: bool = False, per_molecule: bool = True) -> dict:
        """propagate() on inputs already in device buffers (``inp``: x, y, z, vx, vy, vz[, t] DeviceArrays
        of at least n entries; ``state`` a device uint8 array or None).  ``per_molecule=False`` asks
        for no per-molecule label and status (the result then has neither; the work counters still
        count out-of-table and step-capped molecules)."""
        ctx = self.ctx
        a0d, st = a0, state
        res = {"n": n, "labels": list(self.cb.labels)}
        nl, ns = self.cb.n_labels, self.cb.n_states
        tally = self._buf("tally", ns * nl, np.int64).fill(0)
        first = self._buf("first", nl, np.int64).fill(np.iinfo(np.int64).max)
        work = self._buf("work", N.WORK_N, np.uint64).fill(0)
        label = self._buf("label", n, np.int32) if per_molecule else None
        status = self._buf("status", n, np.uint8) if per_molecule else None
        out = {k: self._buf("out_" + k, n) for k in FIELDS} if outputs else None
        out_a = [self._buf(f"out_a{k}", n) for k in range(3)] if want_accel else None
        mask = 0
        for name in aoi_labels:
            if name in self.cb.labels:
                mask |= 1 << self.cb.labels.index(name)
        aoi = {}
        if mask:
            aoi["count"] = self._buf("aoi_count", 1, np.int64).fill(0)
            aoi["index"] = self._buf("aoi_index", n, np.int64)
            aoi["label"] = self._buf("aoi_label", n, np.int32)
            aoi["out"] = {k: self._buf("aoi_out_" + k, n) for k in FIELDS}
        if n:
            self.launch(inp, out=out, out_a=out_a, label=label, status=status, tally=tally, first_index=first,
                        aoi_mask=mask, aoi_cap=n if mask else 0, aoi_count=aoi.get("count"),
                        aoi_index=aoi.get("index"), aoi_label=aoi.get("label"), aoi_out=aoi.get("out"),
                        work=work, a0=a0d, state=st,
                        n=n, index_base=index_base, block=block, flags=flags, refill=refill, noalloc=noalloc)
        ctx.sync()
        res["tally"] = tally.to_host(count=ns * nl).reshape(ns, nl)
        res["first_index"] = first.to_host(count=nl)
        res["work"] = work.to_host(count=N.WORK_N)
        if per_molecule:
            res["label"] = label.to_host(count=n)
            res["status"] = status.to_host(count=n)
        if outputs:
            res["out"] = {k: v.to_host(count=n) for k, v in out.items()}
        if want_accel:
            res["out_a"] = [v.to_host(count=n) for v in out_a]
        if mask:
            cnt = int(aoi["count"].to_host(count=1)[0])
            res["aoi_index"] = aoi["index"].to_host(count=cnt)
            res["aoi_label"] = aoi["label"].to_host(count=cnt)
            res["aoi_out"] = {k: v.to_host(count=cnt) for k, v in aoi["out"].items()}
        return res

    def _record(self, mode, x, y, z, vx, vy, vz, t, a0, state, cap, z_start=0.0, z_step=0.0):
        ctx = self.ctx
        n, inp, a0d, st = self._upload_inputs(x, y, z, vx, vy, vz, t, a0, state)
        nf = 10 if mode == N.REC_ROWS else 6
        buf = ctx.alloc(max(nf * cap * n, 1))
        nrows = ctx.alloc(max(n, 1), np.int32)
        lab = ctx.alloc(max(n, 1), np.int32)
        if n:
            bind(ctx, self.cb)
            Rc = N.Record()
            Rc.n = n
            Rc.index = None
            Rc.inp = _soa(inp)
            for k in range(3):
                Rc.a0[k] = N.addr(a0d[k]) if a0d is not None else None
            Rc.state = N.addr(st)
            Rc.mode, Rc.cap = mode, int(cap)
            Rc.z_start, Rc.z_step = float(z_start), float(z_step)
            Rc.buf, Rc.n_rows, Rc.label = buf.ptr, nrows.ptr, lab.ptr
            N.check("ctraj_record_trajectories", N.lib().ctraj_record_trajectories(ctx.handle, C.byref(Rc)),
                    ctx.handle)
        ctx.sync()
        if mode == N.REC_ZGRID_MOL:
            data = buf.to_host()[: nf * cap * n].reshape(n, cap, nf)
        else:
            data = buf.to_host()[: nf * cap * n].reshape(nf, cap, n)
        return data, nrows.to_host()[:n], lab.to_host()[:n]

    def record_rows(self, x, y, z, vx, vy, vz, t=None, a0=None, state=None, cap: int = 64):
        """Every trajectory row per molecule: (rows [n][cap][10], n_rows [n], labels [n]).

        Row fields: x, y, z, vx, vy, vz, ax, ay, az, t.  If a molecule needs more than ``cap`` rows it
        is re-recorded with a larger cap."""
        while True:
            data, nrows, lab = self._record(N.REC_ROWS, x, y, z, vx, vy, vz, t, a0, state, cap)
            need = int(nrows.max()) if nrows.size else 0
            if need <= cap:
                return np.ascontiguousarray(data.transpose(2, 1, 0)), nrows, lab
            cap = need

    def zgrid(self, x, y, z, vx, vy, vz, t=None, a0=None, state=None, *, z_start: float = 0.0,
              z_step: float = 1e-3, n_samples: int = 2501):
        """Samples at z_k = z_start + k z_step by the post-processing rule: [6][n_samples][n]
        (x, y, vx, vy, vz, t; NaN where the molecule does not reach z_k), written count, labels."""
        return self._record(N.REC_ZGRID, x, y, z, vx, vy, vz, t, a0, state, n_samples, z_start, z_step)

    def zgrid_by_molecule(self, x, y, z, vx, vy, vz, t=None, a0=None, state=None, *, z_start: float = 0.0,
                          z_step: float = 1e-3, n_samples: int = 2501):
        """The samples of ``zgrid`` laid out molecule-major, [n][n_samples][6] (each molecule's samples
        contiguous on the device), written count, labels."""
        return self._record(N.REC_ZGRID_MOL, x, y, z, vx, vy, vz, t, a0, state, n_samples, z_start, z_step)

    def draw_gaussian(self, n: int, mean, sigma, seed: int, offset: int = 0) -> list:
        """Device Gaussian source (Philox4x32-10 + Box-Muller): list of DeviceArrays, one per dimension."""
        return draw_gaussian(self.ctx, n, mean, sigma, seed, offset)


def draw_gaussian(ctx: Context, n: int, mean, sigma, seed: int, offset: int = 0, out=None) -> list:
    mean = np.ascontiguousarray(mean, dtype=np.float64)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    nd = mean.size
    if out is None:
        out = [ctx.alloc(max(int(n), 1)) for _ in range(nd)]
    ptrs = (C.c_void_p * nd)(*[N.addr(o) for o in out])
    N.check("ctraj_draw_gaussian",
            N.lib().ctraj_draw_gaussian(ctx.handle, C.c_uint64(seed), C.c_uint64(offset), int(n), nd,
                                        mean.ctypes.data, sigma.ctypes.data, C.cast(ptrs, C.c_void_p)),
            ctx.handle)
    return out


def draw_disc(ctx: Context, n: int, radius: float, z: float, seed: int, offset: int = 0, out=None) -> list:
    """Device uniform-disc source (CeNTREXPositionDistribution.draw's form, Philox4x32-10): [x, y, z]."""
    if out is None:
        out = [ctx.alloc(max(int(n), 1)) for _ in range(3)]
    N.check("ctraj_draw_disc",
            N.lib().ctraj_draw_disc(ctx.handle, C.c_uint64(seed), C.c_uint64(offset), int(n), float(radius), float(z),
                                    N.addr(out[0]), N.addr(out[1]), N.addr(out[2]) if len(out) > 2 else None),
            ctx.handle)
    return out


def draw_uniform_u8(ctx: Context, n: int, n_values: int, seed: int, offset: int = 0, out=None) -> DeviceArray:
    """Device uniform ids in [0, n_values) (lens-state assignment), Philox4x32-10."""
    if out is None:
        out = ctx.alloc(max(int(n), 1), np.uint8)
    N.check("ctraj_draw_uniform_u8",
            N.lib().ctraj_draw_uniform_u8(ctx.handle, C.c_uint64(seed), C.c_uint64(offset), int(n), int(n_values),
                                          out.ptr), ctx.handle)
    return out


def gather(ctx: Context, n: int, index, index_base: int, src_len: int, src: list, dst: list) -> None:
    """dst[f][j] = src[f][index[j] - index_base] on the device (float64 fields, or one uint8 field)."""
    if n == 0:
        return
    if src and src[0] is not None and src[0].dtype == np.uint8:
        N.check("ctraj_gather_u8", N.lib().ctraj_gather_u8(ctx.handle, int(n), N.addr(index), int(index_base),
                                                           int(src_len), N.addr(src[0]), N.addr(dst[0])), ctx.handle)
        return
    nf = len(src)
    sp = (C.c_void_p * nf)(*[N.addr(a) for a in src])
    dp = (C.c_void_p * nf)(*[N.addr(a) for a in dst])
    N.check("ctraj_gather_f64", N.lib().ctraj_gather_f64(ctx.handle, int(n), N.addr(index), int(index_base), int(src_len),
                                                         nf, C.cast(sp, C.c_void_p), C.cast(dp, C.c_void_p)),
            ctx.handle)


def draw_gaussian_at(ctx: Context, index, n: int, mean, sigma, seed: int, out=None) -> list:
    """Device Gaussian source for molecules ``index[0:n]`` (device int64) of stream ``seed``."""
    mean = np.ascontiguousarray(mean, dtype=np.float64)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    nd = mean.size
    if out is None:
        out = [ctx.alloc(max(int(n), 1)) for _ in range(nd)]
    ptrs = (C.c_void_p * nd)(*[N.addr(o) for o in out])
    N.check("ctraj_draw_gaussian_at",
            N.lib().ctraj_draw_gaussian_at(ctx.handle, C.c_uint64(seed), N.addr(index), int(n), nd, mean.ctypes.data,
                                           sigma.ctypes.data, C.cast(ptrs, C.c_void_p)), ctx.handle)
    return out


def draw_uniform_u8_at(ctx: Context, index, n: int, n_values: int, seed: int, out=None) -> DeviceArray:
    if out is None:
        out = ctx.alloc(max(int(n), 1), np.uint8)
    N.check("ctraj_draw_uniform_u8_at",
            N.lib().ctraj_draw_uniform_u8_at(ctx.handle, C.c_uint64(seed), N.addr(index), int(n), int(n_values),
                                             N.addr(out)), ctx.handle)
    return out


def sort_i64(ctx: Context, keys, n: int) -> None:
    """Ascending in-place sort of the first n device int64 keys."""
    N.check("ctraj_sort_i64", N.lib().ctraj_sort_i64(ctx.handle, N.addr(keys), int(n)), ctx.handle)
