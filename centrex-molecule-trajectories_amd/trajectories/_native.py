"""ctypes binding of libctraj.so (the C ABI declared in include/ctraj.h).

The library is built in-tree by ``trajectories.build`` (``hipcc --offload-arch=gfx950``).  Nothing
here falls back to a CPU path: if the library or a GPU is missing, :func:`lib` / :class:`Context`
raise, so a run either executes the HIP kernels or fails.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libctraj.so"

ABI_VERSION = 1

OK, EINVAL, EHIP, ENOMEM, ESTATE = 0, -1, -2, -3, -4
MAX_ELEMS, MAX_LABELS, MAX_STATES = 64, 64, 4

CIRC, RECT, PLATES, LENS = 1, 2, 3, 4
FORCE_RTABLE, FORCE_WSPLINE_HEX, FORCE_WSPLINE_LIN = 0, 1, 2
INTEG_REF, INTEG_RK4 = 0, 1

ST_OOT, ST_CAPPED = 1, 2
WORK_FACES, WORK_STEPS, WORK_OOT, WORK_CAPPED, WORK_FACES_POST, WORK_N = 0, 1, 2, 3, 4, 8
RUN_STATUS_OR, RUN_SIMPLE, RUN_LANES, RUN_PHASE_TIMES, RUN_QUEUE_ORDER = 1, 2, 4, 8, 16
REC_ROWS, REC_ZGRID, REC_ZGRID_MOL = 1, 2, 3
H2D, D2H, D2D = 1, 2, 3

dptr = C.c_void_p


class Elem(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("label_hit", C.c_int32),
        ("label_inside", C.c_int32),
        ("force", C.c_int32),
        ("integrator", C.c_int32),
        ("table", C.c_int32),
        ("n_steps", C.c_int32),
        ("reserved", C.c_int32),
        ("z0", C.c_double),
        ("z1", C.c_double),
        ("p", C.c_double * 8),
    ]


class Phys(C.Structure):
    _fields_ = [
        ("g", C.c_double),
        ("n_labels", C.c_int32),
        ("label_detected", C.c_int32),
        ("n_states", C.c_int32),
        ("max_lens_steps", C.c_int32),
    ]


class SoA(C.Structure):
    _fields_ = [(k, dptr) for k in ("x", "y", "z", "vx", "vy", "vz", "t")]


class Run(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("index_base", C.c_int64),
        ("inp", SoA),
        ("z_const", C.c_double),
        ("t_const", C.c_double),
        ("a0", dptr * 3),
        ("state", dptr),
        ("out", SoA),
        ("out_a", dptr * 3),
        ("label", dptr),
        ("status", dptr),
        ("tally", dptr),
        ("first_index", dptr),
        ("aoi_mask", C.c_uint64),
        ("aoi_cap", C.c_int64),
        ("aoi_count", dptr),
        ("aoi_index", dptr),
        ("aoi_label", dptr),
        ("aoi_out", SoA),
        ("work", dptr),
        ("block", C.c_int32),
        ("grid", C.c_int32),
        ("flags", C.c_int32),
        ("refill", C.c_int32),
    ]


class Record(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("index", dptr),
        ("inp", SoA),
        ("z_const", C.c_double),
        ("t_const", C.c_double),
        ("a0", dptr * 3),
        ("state", dptr),
        ("mode", C.c_int32),
        ("cap", C.c_int32),
        ("z_start", C.c_double),
        ("z_step", C.c_double),
        ("buf", dptr),
        ("n_rows", dptr),
        ("label", dptr),
        ("block", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Sample(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("row_off", dptr),
        ("n_rows", C.c_int64),
        ("rows", dptr),
        ("z", C.c_double),
        ("out", dptr * 6),
        ("valid", dptr),
    ]


class BatchCfg(C.Structure):
    _fields_ = [
        ("elem_begin", C.c_int32),
        ("n_elems", C.c_int32),
        ("n", C.c_int64),
        ("seed", C.c_uint64),
        ("mean", C.c_double * 6),
        ("sigma", C.c_double * 6),
    ]


class BatchTune(C.Structure):
    _fields_ = [
        ("block", C.c_int32),
        ("chunk", C.c_int32),
        ("grid", C.c_int32),
        ("refill", C.c_int32),
        ("use_graph", C.c_int32),
        ("fused", C.c_int32),
    ]


class Smooth(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("x", dptr),
        ("y", dptr),
        ("vx", dptr),
        ("vy", dptr),
        ("eps", dptr),
        ("z_src", C.c_double),
        ("sigma_vz", C.c_double),
        ("softness", C.c_double),
        ("n_cfg", C.c_int32),
        ("n_params", C.c_int32),
        ("params", dptr),
        ("J", dptr),
        ("grad", dptr),
        ("n_capped", dptr),
        ("ckpt_every", C.c_int32),
        ("max_steps", C.c_int32),
        ("block", C.c_int32),
        ("reserved", C.c_int32),
    ]


# (name, restype, argtypes) of every symbol include/ctraj.h declares
SIGNATURES = [
    ("ctraj_abi_version", C.c_int32, []),
    ("ctraj_ctx_create", C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
    ("ctraj_ctx_destroy", C.c_int32, [C.c_void_p]),
    ("ctraj_last_error", C.c_char_p, [C.c_void_p]),
    ("ctraj_device_info", C.c_int32, [C.c_void_p] + [C.POINTER(C.c_int32)] * 4),
    ("ctraj_dev_alloc", C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("ctraj_dev_free", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("ctraj_host_alloc", C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("ctraj_host_free", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("ctraj_memcpy", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32]),
    ("ctraj_fill_u8", C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint8, C.c_uint64]),
    ("ctraj_fill_i64", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64]),
    ("ctraj_sync", C.c_int32, [C.c_void_p]),
    ("ctraj_event_record", C.c_int32, [C.c_void_p, C.c_int32]),
    ("ctraj_event_elapsed", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_float)]),
    ("ctraj_set_beamline", C.c_int32, [C.c_void_p, C.POINTER(Elem), C.c_int32, C.POINTER(Phys)]),
    ("ctraj_set_rtable", C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32]),
    ("ctraj_set_wspline", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    ("ctraj_propagate", C.c_int32, [C.c_void_p, C.POINTER(Run)]),
    ("ctraj_propagate_noalloc", C.c_int32, [C.c_void_p, C.POINTER(Run)]),
    ("ctraj_phase_times", C.c_int32,
     [C.c_void_p, C.c_int32] + [C.POINTER(C.c_float)] * 3 + [C.POINTER(C.c_int32)]),
    ("ctraj_record_trajectories", C.c_int32, [C.c_void_p, C.POINTER(Record)]),
    ("ctraj_sample_rows", C.c_int32, [C.c_void_p, C.POINTER(Sample)]),
    ("ctraj_draw_gaussian", C.c_int32,
     [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("ctraj_batch_create", C.c_int32,
     [C.c_void_p, C.POINTER(Elem), C.c_int32, C.POINTER(BatchCfg), C.c_int32, C.POINTER(Phys),
      C.POINTER(C.c_void_p)]),
    ("ctraj_batch_run", C.c_int32,
     [C.c_void_p, C.POINTER(BatchTune), C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    ("ctraj_batch_destroy", C.c_int32, [C.c_void_p]),
    ("ctraj_batch_split_path", C.c_int32, [C.c_void_p]),
    ("ctraj_draw_uniform_u8", C.c_int32, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int64, C.c_int32, C.c_void_p]),
    ("ctraj_draw_disc", C.c_int32, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int64, C.c_double, C.c_double,
                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    ("ctraj_draw_gaussian_at", C.c_int32,
     [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("ctraj_draw_uniform_u8_at", C.c_int32, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]),
    ("ctraj_sort_i64", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("ctraj_gather_f64", C.c_int32,
     [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]),
    ("ctraj_gather_u8", C.c_int32, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]),
    ("ctraj_smooth_params", C.c_int32, [C.c_void_p, C.POINTER(C.c_int32), C.c_void_p]),
    ("ctraj_smooth_objective", C.c_int32, [C.c_void_p, C.POINTER(Smooth)]),
]

_LIB = None


class NativeError(RuntimeError):
    """A libctraj.so call returned a negative status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load libctraj.so (once).  Raises if it is not built: there is no fallback path."""
    global _LIB
    if _LIB is None:
        path = Path(os.environ.get("CTRAJ_LIB", LIB_PATH))
        if not path.exists():
            raise RuntimeError(
                f"libctraj.so not found at {path}; build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python -m trajectories.build` -- the trajectory engine has no CPU fallback")
        L = C.CDLL(str(path))
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        v = L.ctraj_abi_version()
        if v != ABI_VERSION:
            raise RuntimeError(f"libctraj.so ABI {v} != expected {ABI_VERSION}")
        _LIB = L
    return _LIB


def check(fn: str, rc: int, ctx=None):
    if rc != OK:
        msg = lib().ctraj_last_error(ctx).decode(errors="replace")
        raise NativeError(fn, rc, msg)


def addr(a) -> int | None:
    """Device address of a DeviceArray / torch tensor, or None."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "ptr"):
        return a.ptr
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    raise TypeError(f"not a device buffer: {type(a)!r}")


_NP = {np.float64: 8, np.int64: 8, np.int32: 4, np.uint8: 1, np.uint64: 8}
