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

# FLOP counting rule: add/sub/mul/div/sqrt = 1, FMA = 2, counted from the kernel code; comparisons, min/max
# clamps, conversions and address arithmetic are not counted.
FLOP_PER_FACE = 28  # time to the plane (sub, div: 2) + drift (dt^2, 3 x position 5, 3 x velocity 2, time: 23)
#                     + squared-radius test (x*x + y*y: 3)
FLOP_PER_RK4_STEP = 82  # folded hexapole RK4 step: 4 slope evaluations (r^2*cu 1, fract 1, spline 2 FMA, slope
#                         mul + FMA: 9 each, plus r^2 = 3 at stages 2-4), stage points 16, slope sums and update
#                         18, bore test r^2 3 (z and t are formed once per molecule from the step count)
FLOP_ADJ_STEP = 204  # reverse of one RK4 step: 4 force+derivative evaluations, stage points, adjoint sums
FLOP_SMOOTH_FACE = 23  # smoothed objective, per aperture face: drift, radius, sigmoid, product
FLOP_SMOOTH_FACE_REV = 34  # its reverse: radius, sigmoid, factor derivative, drift adjoint
PEAK_FP64 = 78.6  # TFLOP/s, MI355X FP64 vector
# The kernels are compute bound on the FP64 vector ALUs ("mfma" in the bench line's bound: compute,
# against the dense FP64 peak); MI355X's FP64 vector peak equals its FP64 matrix peak.
ENGINE = "fp64 vector ALU (MI355X FP64 vector peak = FP64 matrix peak = 78.6 TFLOP/s)"
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
    ap.add_argument("--lanes", action="store_true", help="single-kernel lane-refill walker instead of the two phases")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline wall time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint measurement")
    ap.add_argument("--adjoint-only", action="store_true", help="only the forward+adjoint measurement")
    ap.add_argument("--adjoint-n", type=int, default=1_000_000)
    ap.add_argument("--adjoint-k", type=int, default=0,
                    help="with --adjoint-only: only the forward+adjoint at this checkpoint interval")
    ap.add_argument("--scan-n", type=int, default=100_000, help="molecules per parameter set of the gradient scan")
    ap.add_argument("--scan-sets", type=int, default=64, help="parameter sets of the gradient scan")
    ap.add_argument("--no-extras", action="store_true", help="skip the ballistic and reference-mode measurements")
    ap.add_argument("--extras-only", action="store_true", help="only the ballistic and reference-mode measurements")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scan measurement")
    ap.add_argument("--batch-only", action="store_true", help="only the batched-scan measurement")
    ap.add_argument("--no-stream", action="store_true", help="skip the streamed four-state run")
    ap.add_argument("--stream-only", action="store_true", help="only the streamed four-state run")
    ap.add_argument("--stream-n", type=int, default=100_000_000)
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


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_cores() -> tuple:
    """(cores used, cores in this process's affinity set): every core of the affinity set, within the
    thread budget the host gives a job (OMP_NUM_THREADS) where one is set."""
    aff = len(os.sched_getaffinity(0))
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(aff, cap) if cap > 0 else aff), aff


def cpu_baseline(bl, X, V, seconds: float):
    """pyref.run_simulation_parallel over host cores on the first `sample` molecules."""
    from oracle import pyref

    cores, aff = cpu_cores()
    # time one molecule chunk single-threaded to size the sample
    t0 = time.perf_counter()
    m0 = 200
    pyref.propagate_block(bl.elements, bl.g, X[:, :m0], V[:, :m0])
    per_mol = (time.perf_counter() - t0) / m0
    loops = 100 * cores
    per_chunk = max(1, min(int(seconds * cores / per_mol / loops), X.shape[1] // loops))
    sample = per_chunk * loops
    t0 = time.perf_counter()
    counter, kept = pyref.run_simulation_parallel(bl.elements, bl.g, X[:, :sample], V[:, :sample], ("Detected",),
                                                  n_jobs=cores)
    wall = time.perf_counter() - t0
    assert sum(counter.values()) == sample
    return {"value": sample / wall, "unit": "molecules/s", "cores": cores, "kind": "port",
            "sample": f"first {sample} molecules of the same workload ({loops} joblib chunks of {per_chunk}), "
                      f"per-molecule numpy restatement of run_simulation(n_jobs={cores}); wall {wall:.1f} s; "
                      f"{cores} of the {aff} cores in the affinity set ({cpu_model()}); the restatement skips the "
                      f"reference's per-molecule trajectory preallocation, so it is faster than the reference",
            "cpu_model": cpu_model(), "cores_in_affinity": aff, "counter": counter}


def timed_ms(ctx, fn, reps: int, warmup: int = 2) -> float:
    """Mean device time of fn() over reps back-to-back launches (HIP events on the context stream)."""
    for _ in range(warmup):
        fn()
    ctx.sync()
    ctx.event_record(0)
    for _ in range(reps):
        fn()
    ctx.event_record(1)
    ctx.sync()
    return ctx.event_elapsed_ms(0, 1) / reps


def smooth_flop(theta: np.ndarray, eps: np.ndarray, bl, K: int, max_steps: int, grad: bool) -> float:
    """FLOPs of one k_smooth launch for one parameter set, from each molecule's lens step count: the
    forward sweep, and with grad the reverse sweep's recomputation (per checkpoint segment of length
    Ls: the state every 8 steps, 8 (ceil(Ls/8) - 1) steps, then each 8-step window again, Ls - ceil(Ls/8)
    steps), the adjoint of every step and the faces' reverse."""
    from trajectories import configs

    L, dt = bl.elements[1].L, bl.elements[1].dt
    vz = theta[-1] + configs.VZ_SIGMA * eps
    nfull = np.clip(np.floor(L / (vz * dt)), 0, max_steps)
    S = np.where(vz > 0, nfull + 1, 0).astype(np.int64)
    faces = 4 * int((vz > 0).sum())  # three aperture faces and the lens entrance
    f = FLOP_PER_RK4_STEP * float(S.sum()) + FLOP_SMOOTH_FACE * faces
    if not grad:
        return f
    full, last = S // K, S % K
    def recompute(Ls):
        nsub = -(-Ls // 8)
        return np.where(Ls > 0, 8 * (nsub - 1) + (Ls - nsub), 0)
    rec = full * recompute(np.full_like(S, K)) + recompute(last)
    return f + FLOP_PER_RK4_STEP * float(rec.sum()) + FLOP_ADJ_STEP * float(S.sum()) + FLOP_SMOOTH_FACE_REV * faces


def bench_adjoint(ctx, n: int, K: int, reps: int = 5) -> dict:
    """Forward+adjoint of the smoothed detected flux (config-1 geometry, 10^6 molecules, theta =
    (r1, V0, z0, r2, r3, mu_vz), sigmoid width 50 um, checkpoints every K steps) against the
    2 n_params objective-only evaluations that central differences need."""
    from trajectories import configs
    from trajectories.smooth import SmoothFlux

    bl = configs.config1()
    src = configs.config3_source(n, 3)
    prob = SmoothFlux(bl, *src, sigma_vz=configs.VZ_SIGMA, softness=configs.CONFIG3_SOFTNESS, ckpt_every=K, ctx=ctx)
    theta = configs.config3_theta()
    npar = theta.size
    ms = prob.steps_needed(theta)
    P1 = ctx.to_device(theta)
    J1 = ctx.alloc(1)
    G1 = ctx.alloc(npar)
    tp, tm = theta + 1e-6 * np.abs(theta), theta - 1e-6 * np.abs(theta)
    Pfd = np.repeat(theta[None], 2 * npar, axis=0)
    for j in range(npar):
        Pfd[2 * j, j], Pfd[2 * j + 1, j] = tp[j], tm[j]
    dPfd = ctx.to_device(Pfd.ravel())
    Jfd = ctx.alloc(2 * npar)
    Psep = [ctx.to_device(Pfd[k]) for k in range(2 * npar)]
    Jsep = ctx.alloc(2 * npar)

    t_adj = timed_ms(ctx, lambda: prob.launch(P1.ptr, 1, J1.ptr, G1.ptr, ms), reps)
    t_fwd = timed_ms(ctx, lambda: prob.launch(P1.ptr, 1, J1.ptr, None, ms), reps)
    t_fd_b = timed_ms(ctx, lambda: prob.launch(dPfd.ptr, 2 * npar, Jfd.ptr, None, ms), reps)

    def fd_separate():
        for k in range(2 * npar):
            prob.launch(Psep[k].ptr, 1, Jsep.ptr + 8 * k, None, ms)

    t_fd_s = timed_ms(ctx, fd_separate, max(1, reps // 2), warmup=1)
    grad = G1.to_host()
    Jf = Jfd.to_host()
    fd = (Jf[0::2] - Jf[1::2]) / (tp - tm)
    gs, fs = grad * theta, fd * theta
    big = np.abs(fs) > 1e-8 * np.max(np.abs(fs))
    fd_err = float(np.max(np.abs(gs[big] - fs[big]) / np.abs(fs[big])))
    f_adj = smooth_flop(theta, src[4], bl, K, ms, True)
    f_fwd = smooth_flop(theta, src[4], bl, K, ms, False)
    achieved = f_adj / (t_adj * 1e-3) / 1e12
    return {
        "workload": f"smoothed detected flux over the ES-lens beamline, theta=(r1,V0,z0,r2,r3,mu_vz), sigmoid 50 um, "
                    f"RK4 dt=10 us, checkpoints every {K} steps, source RandomState(3)",
        "molecules": n, "n_params": npar, "ckpt_every": K,
        "fwd_adjoint_ms": t_adj, "fwd_adjoint_molecules_per_s": n / (t_adj * 1e-3),
        "fwd_only_ms": t_fwd, "fwd_only_molecules_per_s": n / (t_fwd * 1e-3), "adjoint_over_fwd": t_adj / t_fwd,
        "fwd_only_tflops": f_fwd / (t_fwd * 1e-3) / 1e12,
        "fd_batched_ms": t_fd_b, "fd_separate_ms": t_fd_s, "fd_evaluations": 2 * npar,
        "speedup_vs_fd_separate": t_fd_s / t_adj, "speedup_vs_fd_batched": t_fd_b / t_adj,
        "grad_vs_fd_max_rel_err": fd_err, "J": float(J1.to_host()[0]),
        "roofline": {"bound": "mfma", "engine": ENGINE, "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "kernel": "k_smooth<true>", "flop_per_launch": f_adj,
                     "flop_rule": "forward %d/RK4 step + 23/face; reverse: recomputed steps %d each (two-level "
                                  "checkpoints), 204 per step adjoint, 34 per face"
                                  % (FLOP_PER_RK4_STEP, FLOP_PER_RK4_STEP)},
    }


def bench_adjoint_scan(ctx, n: int, n_sets: int = 64, K: int = 64, reps: int = 3) -> dict:
    """Batched forward+adjoint scan: n_sets parameter sets (V0, lens z0 and detector radius from the
    batched scan's RandomState(2) draws; r1, r2, mu_vz at their config-3 values) over the same n
    molecules, J and the full gradient of every set in one launch, against the central differences of
    every set (12 objective evaluations per set) as one batched launch and as one launch per set."""
    from trajectories import configs
    from trajectories.smooth import SmoothFlux

    bl = configs.config1()
    src = configs.config3_source(n, 3)
    prob = SmoothFlux(bl, *src, sigma_vz=configs.VZ_SIGMA, softness=configs.CONFIG3_SOFTNESS, ckpt_every=K, ctx=ctx)
    V0, z0, r3, _ = configs.config2_params(n_sets, 2)
    base = configs.config3_theta()
    Th = np.repeat(base[None], n_sets, axis=0)
    Th[:, 1], Th[:, 2], Th[:, 4] = V0, z0, r3
    npar = base.size
    ms = prob.steps_needed(Th)
    dTh = ctx.to_device(Th.ravel())
    J = ctx.alloc(n_sets)
    G = ctx.alloc(n_sets * npar)
    hstep = 1e-6 * np.abs(Th)
    Pfd = np.repeat(Th, 2 * npar, axis=0).reshape(n_sets, 2 * npar, npar)
    for j in range(npar):
        Pfd[:, 2 * j, j] += hstep[:, j]
        Pfd[:, 2 * j + 1, j] -= hstep[:, j]
    dPfd = ctx.to_device(Pfd.ravel())
    Jfd = ctx.alloc(n_sets * 2 * npar)
    t_adj = timed_ms(ctx, lambda: prob.launch(dTh.ptr, n_sets, J.ptr, G.ptr, ms), reps, warmup=1)
    t_fd = timed_ms(ctx, lambda: prob.launch(dPfd.ptr, n_sets * 2 * npar, Jfd.ptr, None, ms), reps, warmup=1)

    def fd_per_set():
        for c in range(n_sets):
            prob.launch(dPfd.ptr + 8 * c * 2 * npar * npar, 2 * npar, Jfd.ptr + 8 * c * 2 * npar, None, ms)

    t_fd_sets = timed_ms(ctx, fd_per_set, 1, warmup=1)
    grad = G.to_host().reshape(n_sets, npar)
    Jf = Jfd.to_host().reshape(n_sets, 2 * npar)
    tp = Pfd[:, 0::2, :][:, np.arange(npar), np.arange(npar)]
    tm = Pfd[:, 1::2, :][:, np.arange(npar), np.arange(npar)]
    fd = (Jf[:, 0::2] - Jf[:, 1::2]) / (tp - tm)
    # gradients scaled by theta (d J / d log theta); each set's error relative to its largest component
    gs, fs = grad * Th, fd * Th
    scale = np.max(np.abs(fs), axis=1, keepdims=True)
    rel = np.abs(gs - fs) / scale
    big = np.abs(fs) > 1e-3 * scale
    rel_comp = np.where(big, np.abs(gs - fs) / np.where(big, np.abs(fs), 1.0), 0.0)
    f_adj = sum(smooth_flop(Th[c], src[4], bl, K, ms, True) for c in range(n_sets))
    achieved = f_adj / (t_adj * 1e-3) / 1e12
    return {
        "workload": f"{n_sets} parameter sets (V0~U[10,40] kV, lens z0~U[0.7,0.9] m, detector r~U[2,5] mm from "
                    f"RandomState(2)) x {n} molecules of the config-3 source, smoothed detected flux and its "
                    f"gradient in theta=(r1,V0,z0,r2,r3,mu_vz), checkpoints every {K} steps",
        "parameter_sets": n_sets, "molecules_per_set": n, "ckpt_every": K,
        "fwd_adjoint_ms": t_adj, "fd_batched_ms": t_fd, "fd_per_set_launches_ms": t_fd_sets,
        "fd_evaluations": n_sets * 2 * npar,
        "speedup_vs_fd_batched": t_fd / t_adj, "speedup_vs_fd_per_set": t_fd_sets / t_adj,
        "grad_vs_fd_max_rel_err": float(rel.max()), "sets_within_1e-5": int((rel.max(axis=1) <= 1e-5).sum()),
        "grad_vs_fd_rel_err_rule": "per set max_j |g_j - fd_j| theta_j / max_j |fd_j theta_j| (central differences, "
                                   "step 1e-6 relative)",
        "grad_vs_fd_max_componentwise_err": float(rel_comp.max()),
        "componentwise_rule": "|g_j - fd_j| / |fd_j| over components above 1e-3 of the set's largest",
        "roofline": {"bound": "mfma", "engine": ENGINE, "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "kernel": "k_smooth<true>", "flop_per_launch": f_adj},
    }


def bench_batch(ctx, reps: int = 5, tune_reps: int = 2) -> dict:
    """The batched scan: 256 ES-lens configurations (V0, lens z0, detector radius, molecule count
    drawn from RandomState(2)), molecules drawn on the device, one persistent graph-replayed launch.
    The autotuner picks the path (two phases or one fused kernel) and its (block, chunk, refill); the
    same configurations run one by one (device draw + ctraj_propagate each) give the unbatched time."""
    from trajectories import configs
    from trajectories.batch import BatchScan
    from trajectories.engine import Engine

    bls, n, seeds = configs.config2()
    scan = BatchScan(bls, n, seeds, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, ctx=ctx)
    t0 = time.perf_counter()
    best = scan.autotune(blocks=(256, 512, 640, 1024), chunks=(128, 256), refills=(8, 12, 16, 24, 32), fused=(0, 1),
                         reps=tune_reps)
    t_tune = time.perf_counter() - t0
    scan.run(best)
    runs = [scan.run(best) for _ in range(reps)]
    ms = float(np.median([r[2] for r in runs]))
    tally, work = runs[0][0], runs[0][1]
    for r in runs[1:]:
        assert np.array_equal(r[0], tally)
    frac = scan.detected_fraction(tally)
    flop = FLOP_PER_FACE * float(work[0]) + FLOP_PER_RK4_STEP * float(work[1])
    achieved = flop / (ms * 1e-3) / 1e12
    # unbatched: the same configurations one at a time on the same device
    engines = [Engine(bl, ctx) for bl in bls]
    ctx.sync()
    wall0 = time.perf_counter()
    ctx.event_record(0)
    for c, eng in enumerate(engines):
        xs = eng.draw_gaussian(int(n[c]), configs.SOURCE_MEAN, configs.SOURCE_SIGMA, int(seeds[c]))
        inp = dict(zip(("x", "y", "z", "vx", "vy", "vz"), xs))
        tal = ctx.alloc(eng.cb.n_labels, np.int64).fill(0)
        eng.launch(inp, tally=tal, n=int(n[c]))
    ctx.event_record(1)
    ctx.sync()
    seq_ms = ctx.event_elapsed_ms(0, 1)
    seq_wall = (time.perf_counter() - wall0) * 1e3
    top = [(b, c, r, f, t) for t, b, c, r, f in scan.tuning[:5]]
    return {
        "workload": "256 ES-lens configurations: V0~U[10,40] kV, lens z0~U[0.7,0.9] m, detector r~U[2,5] mm, "
                    "n~round(10^U[3,6]) from RandomState(2); device Philox source per configuration",
        "configurations": len(bls), "molecules": int(n.sum()),
        "ms": ms, "molecules_per_s": float(n.sum()) / (ms * 1e-3),
        "tune": {"block": best.block, "chunk": best.chunk, "refill": best.refill, "fused": bool(best.fused),
                 "graph": bool(best.use_graph), "candidates": len(scan.tuning), "tuning_s": t_tune,
                 "top5_block_chunk_refill_fused_ms": top},
        "unbatched_device_ms": seq_ms, "unbatched_wall_ms": seq_wall, "speedup_vs_unbatched": seq_ms / ms,
        "detected_fraction": {"min": float(frac.min()), "median": float(np.median(frac)), "max": float(frac.max())},
        "roofline": {"bound": "mfma", "engine": ENGINE, "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64,
                     "kernel": "k_batch_lanes" if best.fused else "k_batch_pre + k_batch_lens", "flop_per_launch": flop},
    }


def bench_stream(ctx, n: int, chunk: int = 1 << 22) -> dict:
    """The streamed run: n molecules of the ES-lens beamline in `chunk`-molecule chunks drawn on the
    device, four lens states assigned uniformly, per-state tallies, and the 1-mm z-grid samples of
    every detected molecule written to one compacted device buffer."""
    from trajectories import configs
    from trajectories.stream import StreamedRun

    run = StreamedRun(configs.config4(), n, chunk=chunk, seed=4, ctx=ctx)
    # one untimed run first: kernel code objects load and the context's sort buffers grow on first use
    warm = run.run(trajectories=True)
    del warm
    t0 = time.perf_counter()
    res = run.run(trajectories=True)
    wall = time.perf_counter() - t0
    det = res.tally[:, res.labels.index("Detected")]
    tb = 6 * 8 * res.n_samples * res.n_detected
    p2 = res.timings["pass2_device_ms"]
    steps = float(res.work[1])
    return {
        "workload": "ES-lens beamline, 4 Stark states (signs -/+, dipole x0.5, x0.8) assigned uniformly, "
                    f"2^{chunk.bit_length() - 1}-molecule chunks drawn on the device (seed 4), detected trajectories "
                    "every 1 mm in z",
        "molecules": n, "chunk": chunk, "chunks": (n + chunk - 1) // chunk,
        "pass1_device_ms": res.timings["pass1_device_ms"], "pass1_wall_s": res.timings["pass1_wall_s"],
        "pass1_molecules_per_s": n / (res.timings["pass1_device_ms"] * 1e-3),
        "pass2_device_ms": p2, "pass2_alloc_s": res.timings["pass2_alloc_s"], "detected": res.n_detected, "detected_per_state": det.tolist(),
        "samples_per_molecule": res.n_samples, "trajectory_bytes": tb,
        "trajectory_write_GBps": tb / (p2 * 1e-3) / 1e9 if p2 > 0 else None,
        "total_wall_s": wall, "rk4_steps_pass1": steps,
    }


def bench_ballistic(ctx, n: int = 100_000_000, reps: int = 10) -> dict:
    """The ballistic beamline (no lens) over n molecules drawn on the device (Philox source, seed 0):
    per-element tallies and the compacted detected phase space; HBM roofline on the algorithmic bytes
    (48 B of inputs per molecule read once, 64 B per kept molecule written once)."""
    from trajectories import _native as N
    from trajectories import configs
    from trajectories.engine import Engine

    eng = Engine(configs.config0(), ctx)
    xs = eng.draw_gaussian(n, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, 0)
    inp = dict(zip(("x", "y", "z", "vx", "vy", "vz"), xs))
    tally = ctx.alloc(eng.cb.n_labels, np.int64).fill(0)
    det = eng.cb.label_detected
    cap = n // 10 + 1024
    cnt = ctx.alloc(1, np.int64).fill(0)
    idx = ctx.alloc(cap, np.int64)
    out = {k: ctx.alloc(cap) for k in ("x", "y", "z", "vx", "vy", "vz", "t")}

    def step():
        cnt.fill(0)
        eng.launch(inp, tally=tally, aoi_mask=1 << det, aoi_cap=cap, aoi_count=cnt, aoi_index=idx, aoi_out=out, n=n)

    for _ in range(2):
        step()
    ctx.sync()
    tally.fill(0)
    ctx.phase_times()
    for _ in range(reps):
        cnt.fill(0)
        eng.launch(inp, tally=tally, aoi_mask=1 << det, aoi_cap=cap, aoi_count=cnt, aoi_index=idx, aoi_out=out, n=n,
                   flags=N.RUN_PHASE_TIMES)
    w, ln, po = ctx.phase_times()
    ms = float(np.mean(w + ln + po))  # one kernel: its whole time is reported as the lens phase
    tal = tally.to_host() // reps
    assert int(tal.sum()) == n
    n_det = int(cnt.to_host()[0])
    nbytes = 48.0 * n + 64.0 * n_det
    gbps = nbytes / (ms * 1e-3) / 1e9
    return {
        "workload": "ballistic beamline: apertures r=6/4/2.5 mm @0.3/0.8/1.2 m, detector r=5 mm @2.0 m; g=9.81; "
                    "source drawn on the device (Philox4x32-10, seed 0)",
        "molecules": n, "ms": ms, "molecules_per_s": n / (ms * 1e-3), "detected": n_det,
        "counts": dict(zip(eng.labels, tal.tolist())),
        "roofline": {"bound": "hbm", "achieved": gbps, "peak": PEAK_HBM, "unit": "GB/s", "frac": gbps / PEAK_HBM,
                     "kernel": "k_propagate<circular apertures, no lens>", "algorithmic_bytes_per_launch": nbytes,
                     "timing": "HIP events recorded by ctraj_propagate around its kernel, mean over %d launches" % reps},
    }


FLOP_PER_REF_STEP = 142  # reference-mode lens step: 4 a(r) evaluations (radius 4, table 5, components 5: 14
#                          each), the reference's stage arithmetic on x, y, z and velocities, time, bore radius


def bench_drop_in(ctx, n: int = 10_000_000) -> dict:
    """The drop-in TrajectorySimulator.run_simulation end to end (wall clock, one call each after a
    small warm-up) on the headline workload: molecules drawn on the host by numpy's global RNG in the
    reference's order (then uploaded), against drawn on the device (device_seed)."""
    from trajectories import configs
    from trajectories.trajectory_simulator import TrajectorySimulator

    bl = configs.config1()
    xdist, vdist = configs.source()
    keep = ["Detected"]
    TrajectorySimulator().run_simulation(bl, "warm", vdist=vdist, xdist=xdist, N_traj=100_000,
                                         apertures_of_interest=keep, ctx=ctx, device_seed=0)
    np.random.seed(1)
    t0 = time.time()
    host = TrajectorySimulator()
    host.run_simulation(bl, "host", vdist=vdist, xdist=xdist, N_traj=n, apertures_of_interest=keep, ctx=ctx)
    t_host = time.time() - t0
    t0 = time.time()
    dev = TrajectorySimulator()
    dev.run_simulation(bl, "dev", vdist=vdist, xdist=xdist, N_traj=n, apertures_of_interest=keep, ctx=ctx,
                       device_seed=1)
    t_dev = time.time() - t0
    return {"workload": "TrajectorySimulator.run_simulation, ES-lens beamline, benchmark source, "
                        "apertures_of_interest=['Detected'], wall clock of one call",
            "molecules": n, "host_draw_s": t_host, "host_draw_molecules_per_s": n / t_host,
            "device_seed_s": t_dev, "device_seed_molecules_per_s": n / t_dev,
            "detected_host_draw": host.counter.counter_dict.get("Detected", 0),
            "detected_device_seed": dev.counter.counter_dict.get("Detected", 0)}


def bench_reference_mode(ctx, n: int = 1_000_000, reps: int = 10, cpu_seconds: float = 10.0, no_cpu: bool = False):
    """The reference's example beamline (examples/lens_simulation_beamline.py:18-72) with its own lens
    model: integrator 'ref' (N = rint(L/dz) steps of dt = dz/vz), linear-field a(r) table; the
    reference's default CeNTREX source replayed with numpy seed 5; per-element tallies and the detected
    molecules.  Beside it the per-molecule CPU restatement of run_simulation on host cores."""
    from trajectories import _native as N
    from trajectories import configs
    from trajectories.distributions import CeNTREXPositionDistribution, CeNTREXVelocityDistribution
    from trajectories.engine import Engine

    bl = configs.reference_example()
    np.random.seed(5)
    vd, xd = CeNTREXVelocityDistribution(), CeNTREXPositionDistribution()
    loops, m = 100, n // 100
    V = np.concatenate([np.asarray(vd.draw(m)).reshape(3, m) for _ in range(loops)], axis=1)
    X = np.concatenate([np.asarray(xd.draw(m)).reshape(3, m) for _ in range(loops)], axis=1)
    eng = Engine(bl, ctx)
    inp = {k: ctx.to_device(a) for k, a in zip(("x", "y", "z", "vx", "vy", "vz"), (X[0], X[1], X[2], V[0], V[1], V[2]))}
    tally = ctx.alloc(eng.cb.n_labels, np.int64).fill(0)
    work = ctx.alloc(N.WORK_N, np.uint64).fill(0)
    det = eng.cb.label_detected
    cap = n // 10 + 1024
    cnt = ctx.alloc(1, np.int64).fill(0)
    idx = ctx.alloc(cap, np.int64)
    out = {k: ctx.alloc(cap) for k in ("x", "y", "z", "vx", "vy", "vz", "t")}
    kw = dict(tally=tally, work=work, aoi_mask=1 << det, aoi_cap=cap, aoi_count=cnt, aoi_index=idx, aoi_out=out, n=n)
    for _ in range(2):
        cnt.fill(0)
        eng.launch(inp, **kw)
    ctx.sync()
    tally.fill(0)
    work.fill(0)
    ctx.phase_times()
    for _ in range(reps):
        cnt.fill(0)
        eng.launch(inp, flags=N.RUN_PHASE_TIMES, **kw)
    w, ln, po = ctx.phase_times()
    tot = w + ln + po
    ms = float(np.mean(tot))
    tal = tally.to_host() // reps
    wk = work.to_host() // reps
    assert int(tal.sum()) == n
    n_oot = int(wk[N.WORK_OOT])
    faces, steps_ = float(wk[N.WORK_FACES]), float(wk[N.WORK_STEPS])
    flop = FLOP_PER_FACE * faces + FLOP_PER_REF_STEP * steps_
    lens_ms = float(np.mean(ln))
    res = {
        "workload": "reference example beamline (4K/40K shields, BB exit, ES lens L=0.6 m in the reference's model: "
                    "integrator 'ref', linear-field a(r) table from the two-level Stark form, field plates, DR "
                    "aperture); CeNTREX default source, numpy seed 5",
        "molecules": n, "ms": ms, "molecules_per_s": n / (ms * 1e-3), "counts": dict(zip(eng.labels, tal.tolist())),
        "out_of_table_molecules": n_oot, "faces": faces, "ref_steps": steps_,
        "roofline": {"bound": "mfma", "engine": ENGINE, "achieved": flop / (ms * 1e-3) / 1e12, "peak": PEAK_FP64,
                     "unit": "TFLOP/s", "frac": flop / (ms * 1e-3) / 1e12 / PEAK_FP64, "flop_per_launch": flop,
                     "flop_rule": "%d per element face, %d per reference lens step" % (FLOP_PER_FACE, FLOP_PER_REF_STEP),
                     "kernels": "PRE + k_lens<RTABLE,REF> + POST" if lens_ms > 0 else "one kernel",
                     "lens_kernel_ms": lens_ms},
    }
    if not no_cpu:
        cb = cpu_baseline(bl, X, V, cpu_seconds)
        res["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        res["gpu_over_cpu"] = res["molecules_per_s"] / cb["value"]
    return res


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
    if args.adjoint_only:
        if args.adjoint_k:
            print(json.dumps({f"K{args.adjoint_k}": bench_adjoint(ctx, args.adjoint_n, args.adjoint_k,
                                                                  max(1, args.steps // 4))}), flush=True)
            return
        print(json.dumps({"K64": bench_adjoint(ctx, args.adjoint_n, 64, max(1, args.steps // 4)),
                          "K8": bench_adjoint(ctx, args.adjoint_n, 8, max(1, args.steps // 4)),
                          "scan": bench_adjoint_scan(ctx, args.scan_n, args.scan_sets)}), flush=True)
        return
    if args.extras_only:
        print(json.dumps({"ballistic": bench_ballistic(ctx),
                          "reference_mode": bench_reference_mode(ctx, no_cpu=args.no_cpu),
                          "run_simulation": bench_drop_in(ctx)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps(bench_batch(ctx)), flush=True)
        return
    if args.stream_only:
        print(json.dumps(bench_stream(ctx, args.stream_n)), flush=True)
        return
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

    flags = N.RUN_SIMPLE if args.simple else (N.RUN_LANES if args.lanes else 0)
    for _ in range(args.warmup):
        step()
    ctx.sync()
    tally.fill(0)
    work.fill(0)
    ctx.sync()
    ctx.phase_times()  # forget anything timed before
    K = args.steps
    kernel_ms, walk_ms, lens_ms, post_ms = [], [], [], []
    tflags = flags | N.RUN_PHASE_TIMES  # the library brackets each kernel of the launch with events on its stream
    wall0 = time.perf_counter()
    ctx.event_record(0)
    for k in range(K):
        aoi_count.fill(0)
        s = 2 + 2 * (k % 31)
        ctx.event_record(s)
        eng.launch(inp, tally=tally, work=work, aoi_mask=1 << det, aoi_cap=cap, aoi_count=aoi_count,
                   aoi_index=aoi_index, aoi_out=aoi_out, n=n, block=args.block, flags=tflags, refill=args.refill)
        ctx.event_record(s + 1)
        if k % 31 == 30:
            kernel_ms += [ctx.event_elapsed_ms(2 + 2 * j, 3 + 2 * j) for j in range(31)]
            w, ln, po = ctx.phase_times()
            walk_ms += list(w)
            lens_ms += list(ln)
            post_ms += list(po)
    ctx.event_record(1)
    ctx.sync()
    wall = time.perf_counter() - wall0
    rem = K % 31
    kernel_ms += [ctx.event_elapsed_ms(2 + 2 * j, 3 + 2 * j) for j in range(rem)]
    w, ln, po = ctx.phase_times()
    walk_ms += list(w)
    lens_ms += list(ln)
    post_ms += list(po)
    assert len(lens_ms) == K, (len(lens_ms), K)
    region_ms = ctx.event_elapsed_ms(0, 1)
    tal = tally.to_host()
    wk = work.to_host()
    assert int(tal.sum()) == n * K, (int(tal.sum()), n * K)
    faces = int(wk[N.WORK_FACES]) / K
    faces_lens = int(wk[N.WORK_FACES_POST]) / K
    steps_ = int(wk[N.WORK_STEPS]) / K
    flop = FLOP_PER_FACE * faces + FLOP_PER_RK4_STEP * steps_
    kavg = float(np.mean(kernel_ms))
    walk_avg, lens_avg = float(np.mean(walk_ms)), float(np.mean(lens_ms))
    achieved = flop / (kavg * 1e-3) / 1e12
    ms_per_step = wall * 1e3 / K
    value = n / (wall / K)
    counts = dict(zip(eng.labels, (tal // K).tolist()))
    n_det = int(aoi_count.to_host()[0])

    traffic = None
    lens = args.config == "lens"
    two_phase = lens and not (args.simple or args.lanes)
    tf = ROOT / "profiles" / ("traffic_propagate_lens.json" if lens else "traffic_propagate_ballistic.json")
    if tf.exists() and n == (10_000_000 if lens else 100_000_000):
        try:
            tj = json.loads(tf.read_text())
            traffic = tj.get("k_lens_bytes_per_launch" if two_phase else "bytes_per_launch")
        except Exception:
            traffic = None
    walk_phase = post_phase = None
    if two_phase:
        # the dominant kernel: k_lens (the lens integration only), timed by the library's events
        flop_lens = FLOP_PER_RK4_STEP * steps_
        ach_lens = flop_lens / (lens_avg * 1e-3) / 1e12
        roof = {"bound": "mfma", "engine": ENGINE, "achieved": ach_lens, "peak": PEAK_FP64, "unit": "TFLOP/s",
                "frac": ach_lens / PEAK_FP64, "traffic": traffic}
        # the other two kernels stream HBM: PRE reads 48 B of inputs per molecule and queues 64 B per
        # molecule entering the lens; the lens kernel rewrites each entry in place, counts the
        # molecules the bore stops and lists the slots of those that leave the lens (lens tallies: a
        # hexapole lens, no per-molecule outputs); POST reads the listed entries (64 B + 4 B of list
        # each) and writes each kept molecule (8 B index + 7 fields)
        n_enter = n - counts.get("Aperture 1", 0) - counts.get("Lens entrance", 0)
        # and the claim list: a bucket byte written and read back and a 4-byte slot per queued molecule
        wbytes = 48.0 * n + 64.0 * n_enter + 6.0 * n_enter
        walk_phase = {"kernel": "k_propagate<PRE> (walk to and through the lens entrance, queue the lens molecules) "
                                "+ k_ord_scan + k_ord_scatter (claim list, longest predicted lens time first)",
                      "ms_avg": walk_avg, "bytes_per_launch": wbytes, "molecules_queued": n_enter,
                      "GBps": wbytes / (walk_avg * 1e-3) / 1e9, "hbm_frac": wbytes / (walk_avg * 1e-3) / 1e9 / PEAK_HBM,
                      "faces": faces - faces_lens}
        n_left = n_enter - counts.get("Inside lens", 0)
        pbytes = 68.0 * n_left + 64.0 * n_det
        post_phase = {"kernel": "k_propagate<POST> (elements after the lens, tallies and kept molecules of the "
                                "molecules that leave the lens; k_lens tallies those it stops)", "ms_avg": float(np.mean(post_ms)), "bytes_per_launch": pbytes,
                      "GBps": pbytes / (float(np.mean(post_ms)) * 1e-3) / 1e9,
                      "hbm_frac": pbytes / (float(np.mean(post_ms)) * 1e-3) / 1e9 / PEAK_HBM, "faces": faces_lens}
    elif lens:
        roof = {"bound": "mfma", "engine": ENGINE, "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64, "traffic": traffic}
    else:
        # no lens: a few divisions per element face against 48 B of inputs per molecule, so HBM is the roof;
        # algorithmic bytes = the inputs read once + each kept molecule's index and 7 fields written once
        gbps = (48.0 * n + 64.0 * n_det) / (kavg * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": gbps, "peak": PEAK_HBM, "unit": "GB/s", "frac": gbps / PEAK_HBM,
                "traffic": traffic, "algorithmic_bytes_per_launch": 48.0 * n + 64.0 * n_det,
                "fp64_tflops": achieved, "fp64_frac": achieved / PEAK_FP64}
    line = {
        "metric": ("molecules/sec through ES-lens beamline (FP64), fwd and fwd+adjoint, vs FP64 roofline" if lens
                   else "molecules/sec through ballistic beamline (FP64), vs HBM roofline"),
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
        "roofline": {**roof,
                     "kernel": ("k_propagate<circular apertures, no lens>" if not lens
                                else "k_propagate<WSPLINE_HEX,RK4>" if args.simple
                                else "k_propagate_lanes<WSPLINE_HEX,RK4>" if args.lanes
                                else "k_lens<WSPLINE_HEX,RK4>"),
                     "kernel_ms_avg": lens_avg if two_phase else kavg,
                     "flop_per_launch": (FLOP_PER_RK4_STEP * steps_) if two_phase else flop,
                     "faces_per_launch": 0 if two_phase else faces, "rk4_steps_per_launch": steps_,
                     "flop_rule": "add/sub/mul/div/sqrt 1, FMA 2; %d per RK4 step, %d per element face"
                                  % (FLOP_PER_RK4_STEP, FLOP_PER_FACE),
                     "timing": "HIP events recorded by ctraj_propagate around each kernel on its stream, "
                               "mean over the %d timed launches" % K},
        "propagation": {"ms_avg": kavg, "flop_per_launch": flop, "tflops": achieved, "fp64_frac": achieved / PEAK_FP64,
                        "faces": faces, "rk4_steps": steps_},
        "region_ms_per_step": region_ms / K,
        "counts": counts,
        "detected": n_det,
        "host_draw_s": t_draw,
    }
    if walk_phase:
        line["walk_phase"] = walk_phase
        line["post_phase"] = post_phase
    if args.config == "lens" and not args.no_adjoint:
        line["fwd_adjoint"] = bench_adjoint(ctx, args.adjoint_n, configs.CONFIG3_CKPT)
        line["fwd_adjoint_K8"] = bench_adjoint(ctx, args.adjoint_n, 8)
        line["fwd_adjoint_scan"] = bench_adjoint_scan(ctx, args.scan_n, args.scan_sets, configs.CONFIG3_CKPT)
    if args.config == "lens" and not args.no_batch:
        line["batch_scan"] = bench_batch(ctx)
    if args.config == "lens" and not args.no_stream:
        line["streamed"] = bench_stream(ctx, args.stream_n, 1 << 22)
        line["streamed_2e25"] = bench_stream(ctx, args.stream_n, 1 << 25)
    if args.config == "lens" and not args.no_extras:
        line["ballistic"] = bench_ballistic(ctx)
        line["reference_mode"] = bench_reference_mode(ctx, cpu_seconds=args.cpu_seconds * 2 / 3, no_cpu=args.no_cpu)
        line["run_simulation"] = bench_drop_in(ctx)
    if not args.no_cpu:
        cb = cpu_baseline(bl, X, V, args.cpu_seconds)
        line["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample", "cpu_model",
                                                   "cores_in_affinity")}
        line["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
