"""GPU: ctraj_propagate_noalloc, the propagation that allocates no device memory, against ctraj_propagate.

Both are called through the C ABI on the same device inputs into separate outputs, twice in a row with
the second run smaller (the scratch array keeps what the first run left there); every per-molecule
output, the tallies, the work counters and the kept molecules must be equal bit for bit.  The lens runs
cover the split kernels over several slices of the scratch queue (with per-molecule outputs, and
without them, where the lens kernel tallies the molecules stopped in the lens), the lane walker
(unfolded lens and CTRAJ_RUN_LANES), reference mode and a beamline without a lens.
"""
import ctypes as C

import numpy as np
import pytest

from trajectories import _native as N
from trajectories import configs
from trajectories.engine import Engine

from beamlines import reference_example_beamline

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "t")


def run_on_device(eng, inp, n, noalloc, flags=0, lean=False):
    """lean: no per-molecule outputs and only the detected molecules kept (the lens kernel then
    tallies the molecules stopped in a hexapole lens itself)."""
    ctx = eng.ctx
    nl, ns = eng.cb.n_labels, eng.cb.n_states
    r = {"tally": ctx.alloc(ns * nl, np.int64).fill(0), "first": ctx.alloc(nl, np.int64).fill(np.iinfo(np.int64).max),
         "work": ctx.alloc(N.WORK_N, np.uint64).fill(0), "count": ctx.alloc(1, np.int64).fill(0),
         "index": ctx.alloc(n, np.int64), "klabel": ctx.alloc(n, np.int32), "kept": {k: ctx.alloc(n) for k in FIELDS}}
    if not lean:
        r.update({"label": ctx.alloc(n, np.int32), "status": ctx.alloc(n, np.uint8),
                  "out": {k: ctx.alloc(n) for k in FIELDS}})
    keep = 1 << eng.cb.label_detected
    if "Inside lens" in eng.labels and not lean:
        keep |= 1 << eng.labels.index("Inside lens")
    eng.launch(inp, out=r.get("out"), label=r.get("label"), status=r.get("status"), tally=r["tally"],
               first_index=r["first"], aoi_mask=keep, aoi_cap=n, aoi_count=r["count"], aoi_index=r["index"],
               aoi_label=r["klabel"], aoi_out=r["kept"], work=r["work"], n=n, flags=flags, noalloc=noalloc)
    ctx.sync()
    cnt = int(r["count"].to_host()[0])
    idx = r["index"].to_host()[:cnt]
    o = np.argsort(idx)
    res = {"tally": r["tally"].to_host(), "first": r["first"].to_host(), "work": r["work"].to_host(), "index": idx[o],
           "klabel": r["klabel"].to_host()[:cnt][o], "kept": {k: v.to_host()[:cnt][o] for k, v in r["kept"].items()}}
    if not lean:
        res.update({"label": r["label"].to_host(), "status": r["status"].to_host(),
                    "out": {k: v.to_host() for k, v in r["out"].items()}})
    return res


def assert_same(a, b, tag):
    for k in ("tally", "first", "label", "status", "index", "klabel"):
        if k in a or k in b:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{tag}: {k}")
    np.testing.assert_array_equal(a["work"][:4], b["work"][:4], err_msg=f"{tag}: work")
    for k in FIELDS:
        if "out" in a or "out" in b:
            np.testing.assert_array_equal(a["out"][k], b["out"][k], err_msg=f"{tag}: out {k}")
        np.testing.assert_array_equal(a["kept"][k], b["kept"][k], err_msg=f"{tag}: kept {k}")


CASES = [
    ("ES lens, split kernels over several scratch slices", lambda: configs.config1(), (2_500_000, 100_003), 0, 61),
    ("ES lens, lane walker", lambda: configs.config1(), (300_000, 1000), N.RUN_LANES, 62),
    ("unfolded lens (V0 = 0), lane walker", lambda: configs.config1(V0=0.0), (200_000, 4321), 0, 63),
    ("reference-mode lens", lambda: reference_example_beamline(), (400_000, 777), 0, 64),
    ("ballistic beamline", lambda: configs.config0(), (1_000_000, 65), 0, 65),
    ("ES lens, split kernels with lens tallies, no per-molecule outputs", lambda: configs.config1(),
     (2_500_000, 100_003), 0, 67),
]


@pytest.mark.parametrize("name,make,sizes,flags,seed", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_propagate_on_two_calls_in_a_row(ctx, name, make, sizes, flags, seed):
    bl = make()
    eng = Engine(bl, ctx)
    for n in sizes:
        X, V = configs.draw_source(-(-n // 100) * 100, seed=seed + n)
        X, V = X[:, :n], V[:, :n]
        if name == "reference-mode lens":
            X[:2] *= 0.05
            V[:2] *= 0.05
        inp = {k: ctx.to_device(np.ascontiguousarray(a)) for k, a in zip(FIELDS[:6], (X[0], X[1], X[2], V[0], V[1], V[2]))}
        lean = "no per-molecule outputs" in name
        ref = run_on_device(eng, inp, n, noalloc=False, flags=flags, lean=lean)
        got = run_on_device(eng, inp, n, noalloc=True, flags=flags, lean=lean)
        assert int(got["tally"].sum()) == n
        assert_same(got, ref, f"{name}, n={n}")


def test_noalloc_refuses_what_would_allocate_or_overflow(ctx):
    eng = Engine(configs.config1(), ctx)
    X, V = configs.draw_source(1000, seed=66)
    inp = {k: ctx.to_device(a) for k, a in zip(FIELDS[:6], (X[0], X[1], X[2], V[0], V[1], V[2]))}
    tally = ctx.alloc(eng.cb.n_labels, np.int64).fill(0)
    with pytest.raises(N.NativeError, match="PHASE_TIMES"):
        eng.launch(inp, tally=tally, n=1000, flags=N.RUN_PHASE_TIMES, noalloc=True)
    with pytest.raises(N.NativeError, match="scratch"):
        eng.launch(inp, tally=tally, n=1000, grid=1 << 20, noalloc=True)
    R = N.Run()
    R.n = 5
    assert N.lib().ctraj_propagate_noalloc(ctx.handle, C.byref(R)) == N.EINVAL  # no inputs
    eng.launch(inp, tally=tally, n=1000, noalloc=True)
    ctx.sync()
    assert int(tally.to_host().sum()) == 1000
