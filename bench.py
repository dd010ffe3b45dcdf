#!/usr/bin/env python3
"""Benchmark: molecules/s through the ES-lens beamline (FP64) on one MI355X.

Workload (default): the ES-lens beamline -- 6 mm aperture at 0.3 m; hexapole lens z0 = 0.8 m,
L = 0.6 m, R = 1.75 cm, V0 = 30 kV, W(|E|) two-level spline; 3 mm aperture at 1.6 m; detector
r = 4 mm at 2.5 m; classical RK4 at dt = 10 us -- over 10^7 molecules of the benchmark source
(seed 1, reference draw order), resident in HBM before timing.  One step = one ctraj_propagate over
all molecules producing per-element tallies and the compacted detected-molecule phase space.

Prints one JSON line.  The CPU baseline is the per-molecule restatement of the reference's
run_simulation(n_jobs = host cores) on a bounded sample of the same molecules.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (str(ROOT), str(ROOT / "centrex-molecule-trajectories_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

FLOP_PER_FACE = 29  # drift to a plane (25) + radius test (4); see DESIGN.md
FLOP_PER_RK4_STEP = 106  # classical RK4 step with 4 spline-force evaluations + bore test
PEAK_FP64 = 78.6  # TFLOP/s, MI355X FP64 vector
PEAK_HBM = 8000.0  # GB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--config", choices=["lens", "ballistic"], default="lens")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0, help="lane-refill threshold (0 = library default)")
    ap.add_argument("--simple", action="store_true", help="one molecule per thread (no lane refill)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline wall time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of known size (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration():
    """A plain 1 GiB device-to-device copy through libctraj (for FETCH_SIZE / WRITE_SIZE scaling)."""
    from trajectories import _native as N
    from trajectories.device import get_context

    ctx = get_context(0)
    n = 1 << 27  # doubles = 1 GiB
    a = ctx.alloc(n).fill(1.0)
    b = ctx.alloc(n)
    ctx.sync()
    for _ in range(3):
        ctx.memcpy(b.ptr, a.ptr, n * 8, N.D2D)
    ctx.sync()
    print(json.dumps({"copy_bytes": n * 8, "reps": 3}))


def cpu_baseline(bl, X, V, seconds: float):
    """pyref.run_simulation_parallel over host cores on the first `sample` molecules."""
    from oracle import pyref

    cores = min(16, len(os.sched_getaffinity(0)))
    # time one molecule chunk single-threaded to size the sample
    t0 = time.perf_counter()
    m0 = 200
    pyref.propagate_block(bl.elements, bl.g, X[:, :m0], V[:, :m0])
    per_mol = (time.perf_counter() - t0) / m0
    loops = 100 * cores
    per_chunk = max(1, int(seconds * cores / per_mol / loops))
    sample = per_chunk * loops
    t0 = time.perf_counter()
    counter, kept = pyref.run_simulation_parallel(bl.elements, bl.g, X[:, :sample], V[:, :sample], ("Detected",),
                                                  n_jobs=cores)
    wall = time.perf_counter() - t0
    assert sum(counter.values()) == sample
    return {"value": sample / wall, "unit": "molecules/s", "cores": cores, "kind": "port",
            "sample": f"first {sample} molecules of the same workload ({loops} joblib chunks of {per_chunk}), "
                      f"per-molecule numpy restatement of run_simulation(n_jobs={cores}); wall {wall:.1f} s",
            "counter": counter}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from trajectories import _native as N
    from trajectories import configs
    from trajectories.device import get_context
    from trajectories.engine import Engine

    ctx = get_context(0)
    if args.config == "lens":
        bl = configs.config1()
        seed = 1
        workload = ("ES-lens beamline: aperture r=6 mm @0.3 m; hexapole lens z0=0.8 m L=0.6 m R=1.75 cm V0=30 kV "
                    "(two-level W(|E|) spline, 1001 knots); aperture r=3 mm @1.6 m; detector r=4 mm @2.5 m; "
                    "classical RK4 dt=10 us; g=9.81")
    else:
        bl = configs.config0()
        seed = 0
        workload = "ballistic beamline: apertures r=6/4/2.5 mm @0.3/0.8/1.2 m, detector r=5 mm @2.0 m; g=9.81"
    n = args.n
    t0 = time.time()
    X, V = configs.draw_source(n, seed)
    t_draw = time.time() - t0
    eng = Engine(bl, ctx)
    inp = {k: ctx.to_device(a) for k, a in zip(("x", "y", "z", "vx", "vy", "vz"), (X[0], X[1], X[2], V[0], V[1], V[2]))}
    nl = eng.cb.n_labels
    tally = ctx.alloc(nl, np.int64).fill(0)
    work = ctx.alloc(N.WORK_N, np.uint64).fill(0)
    det = eng.cb.label_detected
    cap = n // 10 + 1024
    aoi_count = ctx.alloc(1, np.int64).fill(0)
    aoi_index = ctx.alloc(cap, np.int64)
    aoi_out = {k: ctx.alloc(cap) for k in ("x", "y", "z", "vx", "vy", "vz", "t")}

    def step():
        aoi_count.fill(0)
        eng.launch(inp, tally=tally, work=work, aoi_mask=1 << det, aoi_cap=cap, aoi_count=aoi_count,
                   aoi_index=aoi_index, aoi_out=aoi_out, n=n, block=args.block, flags=flags, refill=args.refill)

    flags = N.RUN_SIMPLE if args.simple else 0
    for _ in range(args.warmup):
        step()
    ctx.sync()
    tally.fill(0)
    work.fill(0)
    ctx.sync()
    K = args.steps
    kernel_ms = []
    wall0 = time.perf_counter()
    ctx.event_record(0)
    for k in range(K):
        aoi_count.fill(0)
        s = 2 + 2 * (k % 31)
        ctx.event_record(s)
        eng.launch(inp, tally=tally, work=work, aoi_mask=1 << det, aoi_cap=cap, aoi_count=aoi_count,
                   aoi_index=aoi_index, aoi_out=aoi_out, n=n, block=args.block, flags=flags, refill=args.refill)
        ctx.event_record(s + 1)
        if k % 31 == 30:
            kernel_ms += [ctx.event_elapsed_ms(2 + 2 * j, 3 + 2 * j) for j in range(31)]
    ctx.event_record(1)
    ctx.sync()
    wall = time.perf_counter() - wall0
    rem = K % 31
    kernel_ms += [ctx.event_elapsed_ms(2 + 2 * j, 3 + 2 * j) for j in range(rem)]
    region_ms = ctx.event_elapsed_ms(0, 1)
    tal = tally.to_host()
    wk = work.to_host()
    assert int(tal.sum()) == n * K, (int(tal.sum()), n * K)
    faces = int(wk[N.WORK_FACES]) / K
    steps_ = int(wk[N.WORK_STEPS]) / K
    flop = FLOP_PER_FACE * faces + FLOP_PER_RK4_STEP * steps_
    kavg = float(np.mean(kernel_ms))
    achieved = flop / (kavg * 1e-3) / 1e12
    ms_per_step = wall * 1e3 / K
    value = n / (wall / K)
    counts = dict(zip(eng.labels, (tal // K).tolist()))
    n_det = int(aoi_count.to_host()[0])

    traffic = None
    tf = ROOT / "profiles" / "traffic_propagate_lens.json"
    if args.config == "lens" and tf.exists() and n == 10_000_000:
        try:
            traffic = json.loads(tf.read_text()).get("bytes_per_launch")
        except Exception:
            traffic = None
    line = {
        "metric": "molecules/sec through ES-lens beamline (FP64), fwd and fwd+adjoint, vs FP64 roofline",
        "value": value,
        "unit": "molecules/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (benchmark source drawn with RandomState(%d) in the reference's order)" % seed,
        "config": {"workload": workload, "molecules": n, "outputs": "per-element tallies + compacted detected phase space"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": traffic,
                     "kernel": "k_propagate<WSPLINE_HEX,RK4>" if args.simple or args.config != "lens"
                     else "k_propagate_lanes<WSPLINE_HEX,RK4>", "kernel_ms_avg": kavg,
                     "flop_per_launch": flop, "faces_per_launch": faces, "rk4_steps_per_launch": steps_},
        "region_ms_per_step": region_ms / K,
        "counts": counts,
        "detected": n_det,
        "host_draw_s": t_draw,
    }
    if not args.no_cpu:
        cb = cpu_baseline(bl, X, V, args.cpu_seconds)
        line["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        line["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
