"""ctypes front end of the C oracle (oracle/ctraj_oracle.c).

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product package never imports it.

It consumes the same packed element table as the device (trajectories.compile), so a
compiled beamline can be run through both and compared bit for bit.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libctraj_oracle.so"
FIELDS = ("x", "y", "z", "vx", "vy", "vz", "t")

_L = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _L
    if _L is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp = C.c_void_p
        L.orc_propagate.restype = C.c_int
        L.orc_propagate.argtypes = [vp, C.c_int, vp, vp, C.c_int, vp, vp, vp, C.c_int64, vp, C.c_double,
                                    C.c_double, vp, vp, vp, vp, vp, vp, vp, C.c_int32, vp, vp, vp, C.c_int]
        L.orc_zgrid_from_rows.restype = None
        L.orc_zgrid_from_rows.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, vp]
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [vp, C.c_uint32, C.c_uint32]
        L.orc_draw_gaussian.restype = None
        L.orc_draw_gaussian.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_int, vp, vp, vp, vp]
        L.orc_draw_uniform_u8.restype = None
        L.orc_draw_uniform_u8.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_int, vp]
        L.orc_draw_disc.restype = None
        L.orc_draw_disc.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_double, C.c_double, vp, vp, vp]
        L.orc_smooth.restype = C.c_int
        L.orc_smooth.argtypes = [vp, C.c_int, C.c_double, vp, C.c_int, C.c_int64, vp, vp, vp, vp, vp, C.c_double,
                                 C.c_double, C.c_double, C.c_int, C.c_int, vp, C.c_int, vp, vp, vp, vp, C.c_int]
        _L = L
    return _L


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[None if a is None else a.ctypes.data for a in arrs])


def propagate(cb, x, y, z, vx, vy, vz, t=None, a0=None, state=None, *, rows_cap: int = 0, threads: int | None = None,
              z_const: float = 0.0, t_const: float = 0.0):
    """Run the oracle over host arrays for a trajectories.compile.CompiledBeamline.

    Returns dict(out, out_a, label, status, edge, work[, rows, n_rows])."""
    L = lib()
    ins = [np.ascontiguousarray(v, dtype=np.float64) if v is not None else None for v in (x, y, z, vx, vy, vz, t)]
    n = ins[0].size
    outs = [np.empty(n) for _ in FIELDS]
    outa = [np.empty(n) for _ in range(3)]
    a0s = None
    if a0 is not None:
        a0s = [np.ascontiguousarray(a, dtype=np.float64) for a in a0]
    st = np.ascontiguousarray(state, dtype=np.uint8) if state is not None else None
    label = np.empty(n, np.int32)
    status = np.empty(n, np.uint8)
    edge = np.empty(n)
    work = np.zeros(8, np.uint64)
    rows = np.full((n, rows_cap, 10), np.nan) if rows_cap else None
    n_rows = np.empty(n, np.int32)
    coef = cb.wspline_coef()
    nseg = 0 if coef is None else coef.shape[1]
    rt_r = [None] * 8
    rt_a = [None] * 8
    rt_n = np.zeros(8, np.int32)
    for i, (r, a) in enumerate(cb.rtables):
        rt_r[i], rt_a[i], rt_n[i] = r, a, r.size
    keep = (rt_r, rt_a)  # noqa: F841 (lifetime)
    L.orc_propagate(C.addressof(cb.elems), cb.n_elems, C.addressof(cb.phys),
                    None if coef is None else coef.ctypes.data, nseg, _ptrs(rt_r), _ptrs(rt_a), rt_n.ctypes.data,
                    n, _ptrs(ins), float(z_const), float(t_const), _ptrs(a0s) if a0s else None,
                    None if st is None else st.ctypes.data, _ptrs(outs), _ptrs(outa), label.ctypes.data,
                    status.ctypes.data, None if rows is None else rows.ctypes.data, int(rows_cap),
                    n_rows.ctypes.data, edge.ctypes.data, work.ctypes.data,
                    int(threads if threads is not None else min(16, os.cpu_count() or 1)))
    res = {"out": dict(zip(FIELDS, outs)), "out_a": outa, "label": label, "status": status, "edge": edge,
           "work": work, "n_rows": n_rows}
    if rows is not None:
        res["rows"] = rows
    return res


def zgrid_from_rows(rows: np.ndarray, n_rows: int, z_start: float, z_step: float, cap: int) -> np.ndarray:
    """post_processing.py resampling of one molecule's rows [n][10] -> [6][cap]."""
    r = np.ascontiguousarray(rows[:n_rows], dtype=np.float64)
    out = np.empty((6, cap))
    lib().orc_zgrid_from_rows(r.ctypes.data, int(n_rows), float(z_start), float(z_step), int(cap), out.ctypes.data)
    return out


def philox(counter, key):
    c = np.ascontiguousarray(counter, dtype=np.uint32).copy()
    lib().orc_philox4x32_10(c.ctypes.data, int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF)
    return c


def draw_gaussian(n, mean, sigma, seed, offset=0, raw=False):
    mean = np.ascontiguousarray(mean, dtype=np.float64)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    nd = mean.size
    outs = [np.empty(n) for _ in range(nd)]
    rawbuf = np.empty((n, (nd + 1) // 2, 4), np.uint32) if raw else None
    lib().orc_draw_gaussian(C.c_uint64(seed), C.c_uint64(offset), int(n), nd, mean.ctypes.data, sigma.ctypes.data,
                            _ptrs(outs), None if rawbuf is None else rawbuf.ctypes.data)
    return (outs, rawbuf) if raw else outs


def draw_disc(n, radius, z, seed, offset=0):
    """[x, y, z] of ctraj_draw_disc's uniform-disc source."""
    outs = [np.empty(n) for _ in range(3)]
    lib().orc_draw_disc(C.c_uint64(seed), C.c_uint64(offset), int(n), float(radius), float(z), outs[0].ctypes.data,
                        outs[1].ctypes.data, outs[2].ctypes.data)
    return outs


def draw_uniform_u8(n, n_values, seed, offset=0):
    out = np.empty(n, np.uint8)
    lib().orc_draw_uniform_u8(C.c_uint64(seed), C.c_uint64(offset), int(n), int(n_values), out.ctypes.data)
    return out


def smooth(cb, x, y, vx, vy, eps, params, *, sigma_vz: float, softness: float, z_src: float = 0.0,
           max_steps: int = 4096, grad: bool = True, per_molecule: bool = False, threads: int | None = None):
    """CPU FP64 smoothed detected flux J[c] and its gradient grad[c][j] (reverse mode) for the
    parameter sets params [n_cfg][n_params] (see ctraj_smooth_objective).

    Returns dict(J, grad (or None), n_capped[, per_mol [n_cfg][n][n_params + 1]])."""
    L = lib()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, vx, vy, eps)]
    n = arrs[0].size
    P = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
    n_cfg, n_params = P.shape
    coef = cb.wspline_coef()
    ws = np.zeros((1, 4)) if coef is None else np.ascontiguousarray(coef[0])
    nseg = 0 if coef is None else coef.shape[1]
    J = np.empty(n_cfg)
    G = np.empty((n_cfg, n_params)) if grad else None
    pm = np.empty((n_cfg, n, n_params + 1)) if per_molecule else None
    capped = np.zeros(1, np.int64)
    rc = L.orc_smooth(C.addressof(cb.elems), cb.n_elems, float(cb.phys.g), ws.ctypes.data, nseg, n,
                      *[a.ctypes.data for a in arrs], float(z_src), float(sigma_vz), float(softness), n_cfg,
                      n_params, P.ctypes.data, int(max_steps), J.ctypes.data,
                      None if G is None else G.ctypes.data, None if pm is None else pm.ctypes.data,
                      capped.ctypes.data, int(threads if threads is not None else min(16, os.cpu_count() or 1)))
    if rc != 0:
        raise ValueError("oracle smooth objective supports circular apertures and lenses only")
    res = {"J": J, "grad": G, "n_capped": int(capped[0])}
    if pm is not None:
        res["per_mol"] = pm
    return res
