"""Device context and buffers on top of libctraj.so.

Device memory is owned by :class:`DeviceArray` objects allocated through the library
(``ctraj_dev_alloc``); copies and kernels are queued on the context's single HIP stream.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _native as N


class DeviceArray:
    """A 1-D device buffer of ``n`` elements of ``dtype`` (owned; freed on ``free()``/GC)."""

    __slots__ = ("ctx", "n", "dtype", "nbytes", "ptr", "__weakref__")

    def __init__(self, ctx: "Context", n: int, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        N.check("ctraj_dev_alloc", N.lib().ctraj_dev_alloc(ctx.handle, self.nbytes, C.byref(p)), ctx.handle)
        self.ptr = p.value

    def upload(self, host: np.ndarray, sync: bool = True):
        a = np.ascontiguousarray(host, dtype=self.dtype).reshape(-1)
        if a.size != self.n:
            raise ValueError(f"size mismatch {a.size} != {self.n}")
        self.ctx.memcpy(self.ptr, a.ctypes.data, self.nbytes, N.H2D)
        if sync:
            self.ctx.sync()
        return self

    def to_host(self, out: np.ndarray | None = None, count: int | None = None) -> np.ndarray:
        """Copy the buffer (or its first ``count`` elements) to the host."""
        k = self.n if count is None else max(0, min(int(count), self.n))
        if out is None:
            out = np.empty(k, dtype=self.dtype)
        if k:
            self.ctx.memcpy(out.ctypes.data, self.ptr, k * self.dtype.itemsize, N.D2H)
        self.ctx.sync()
        return out

    def fill(self, value):
        if self.dtype.itemsize == 8:
            bits = np.array([value], dtype=self.dtype).view(np.int64)[0]
            N.check("ctraj_fill_i64", N.lib().ctraj_fill_i64(self.ctx.handle, self.ptr, int(bits), self.n),
                    self.ctx.handle)
        elif self.dtype.itemsize == 1:
            N.check("ctraj_fill_u8", N.lib().ctraj_fill_u8(self.ctx.handle, self.ptr, int(value), self.nbytes),
                    self.ctx.handle)
        else:
            raise TypeError("fill supports 1- and 8-byte dtypes")
        return self

    def free(self):
        if self.ptr and self.ctx.handle:
            N.lib().ctraj_dev_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self):
        return self.n


class Context:
    """One device + one HIP stream (``ctraj_ctx``)."""

    def __init__(self, device: int = 0):
        L = N.lib()
        h = C.c_void_p()
        rc = L.ctraj_ctx_create(int(device), C.byref(h))
        if rc != N.OK:
            raise N.NativeError("ctraj_ctx_create", rc, L.ctraj_last_error(None).decode(errors="replace"))
        self.handle = h.value
        self.device = int(device)
        ncu, lds, wave, arch = (C.c_int32() for _ in range(4))
        N.check("ctraj_device_info",
                L.ctraj_device_info(self.handle, C.byref(ncu), C.byref(lds), C.byref(wave), C.byref(arch)),
                self.handle)
        self.n_cu, self.lds_per_block, self.wave, self.arch = ncu.value, lds.value, wave.value, arch.value
        self.bound = None  # the CompiledBeamline currently set on the device
        self.lock = threading.RLock()

    # --- memory ---
    def alloc(self, n, dtype=np.float64) -> DeviceArray:
        return DeviceArray(self, n, dtype)

    def to_device(self, host, dtype=None) -> DeviceArray:
        host = np.asarray(host)
        d = DeviceArray(self, host.size, dtype or host.dtype)
        return d.upload(host)

    def memcpy(self, dst: int, src: int, nbytes: int, kind: int):
        N.check("ctraj_memcpy", N.lib().ctraj_memcpy(self.handle, dst, src, int(nbytes), int(kind)), self.handle)

    def sync(self):
        N.check("ctraj_sync", N.lib().ctraj_sync(self.handle), self.handle)

    # --- timing on the context stream ---
    def event_record(self, slot: int):
        N.check("ctraj_event_record", N.lib().ctraj_event_record(self.handle, int(slot)), self.handle)

    def event_elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        N.check("ctraj_event_elapsed", N.lib().ctraj_event_elapsed(self.handle, int(a), int(b), C.byref(ms)),
                self.handle)
        return float(ms.value)

    def phase_times(self, max_n: int = 64):
        """Kernel times (ms) of the launches made with RUN_PHASE_TIMES since the last call, oldest first:
        (walk to the lens, lens, walk after the lens) arrays; a single-kernel run has all its time in the
        lens array."""
        bufs = [(C.c_float * max_n)() for _ in range(3)]
        cnt = C.c_int32()
        N.check("ctraj_phase_times", N.lib().ctraj_phase_times(self.handle, int(max_n), *bufs, C.byref(cnt)),
                self.handle)
        k = cnt.value
        return tuple(np.array(b[:k], dtype=np.float64) for b in bufs)

    def close(self):
        if getattr(self, "handle", None):
            N.lib().ctraj_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT: dict[int, Context] = {}
_LOCK = threading.Lock()


def get_context(device: int = 0) -> Context:
    """Process-wide context for ``device`` (created on first use)."""
    with _LOCK:
        ctx = _DEFAULT.get(device)
        if ctx is None or ctx.handle is None:
            ctx = Context(device)
            _DEFAULT[device] = ctx
        return ctx
