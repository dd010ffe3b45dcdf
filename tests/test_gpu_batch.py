"""GPU: the batched scan (ctraj_batch_*) against per-configuration propagation.

Each configuration's molecules are the device source draws (ctraj_draw_gaussian with the
configuration's seed); the same draws are propagated configuration by configuration with
ctraj_propagate and, copied to the host, with the CPU oracle.  Tallies must agree exactly.
Graph replays are checked against eager launches, with eager work between the replays.
"""
import numpy as np
import pytest

from oracle import oracle as O
from trajectories import configs
from trajectories.batch import BatchScan
from trajectories.beamline import Beamline
from trajectories.beamline_elements import CircularAperture
from trajectories.engine import Engine

pytestmark = pytest.mark.gpu


def small_scan(ctx):
    V0 = [30e3, 12e3, 38e3, 25e3, 30e3, 20e3]
    z0 = [0.8, 0.72, 0.88, 0.8, 0.75, 0.8]
    r = [4e-3, 2.2e-3, 4.9e-3, 3e-3, 4e-3, 3.5e-3]
    n = [20_000, 1_000, 7_777, 0, 12_345, 64]
    bls = [configs.config1(V0=V0[c], lens_z0=z0[c], r3=r[c]) for c in range(6)]
    # a configuration without a lens rides along
    bls.append(Beamline([CircularAperture("Aperture 1", 0.3, 0.0, d=12e-3),
                         CircularAperture("Detector", 2.0, 0.0, d=10e-3)], configs.G))
    n.append(5_000)
    seeds = np.arange(len(bls), dtype=np.uint64) + np.uint64(12345)
    return bls, np.array(n), seeds, BatchScan(bls, n, seeds, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, ctx=ctx)


def per_config_reference(ctx, scan, bls, n, seeds):
    """Tallies of ctraj_propagate and of the oracle on each configuration's device draws."""
    dev, orc = [], []
    for c, bl in enumerate(bls):
        eng = Engine(bl, ctx)
        if n[c] == 0:
            dev.append({})
            orc.append({})
            continue
        cols = eng.draw_gaussian(int(n[c]), configs.SOURCE_MEAN, configs.SOURCE_SIGMA, int(seeds[c]))
        x, y, z, vx, vy, vz = [a.to_host() for a in cols]
        r = eng.propagate(x, y, z, vx, vy, vz, outputs=False)
        dev.append(dict(zip(eng.labels, r["tally"].sum(axis=0).tolist())))
        o = O.propagate(eng.cb, x, y, z, vx, vy, vz)
        lab = np.bincount(o["label"], minlength=eng.cb.n_labels)
        orc.append(dict(zip(eng.labels, lab.tolist())))
    return dev, orc


def as_dict(scan, tally, c):
    return {scan.labels[k]: int(tally[c, k]) for k in range(len(scan.labels)) if tally[c, k]}


def test_batch_equals_per_configuration_propagation(ctx):
    bls, n, seeds, scan = small_scan(ctx)
    tally, work, ms = scan.run(graph=False)
    assert tally.sum() == n.sum()
    dev, orc = per_config_reference(ctx, scan, bls, n, seeds)
    for c in range(len(bls)):
        got = as_dict(scan, tally, c)
        want = {k: v for k, v in dev[c].items() if v}
        assert got == want, (c, got, want)
        assert got == {k: v for k, v in orc[c].items() if v}, c
    print("batch ms", ms, "work", work[:4])


def test_graph_replays_match_eager_runs(ctx):
    bls, n, seeds, scan = small_scan(ctx)
    eager, _, _ = scan.run(graph=False)
    for _ in range(3):
        g, _, _ = scan.run(graph=True)
        np.testing.assert_array_equal(g, eager)
        e2, _, _ = scan.run(graph=False)  # eager work between the replays
        np.testing.assert_array_equal(e2, eager)


def test_every_tuning_gives_the_same_tallies(ctx):
    bls, n, seeds, scan = small_scan(ctx)
    ref, _, _ = scan.run(graph=False)
    best = scan.autotune(blocks=(64, 256, 1024), chunks=(32, 256), refills=(1, 16, 64), fused=(0, 1), reps=1)
    assert {f for *_, f in scan.tuning} == {0, 1}
    for _, b, c, r, f in scan.tuning:
        t, w, _ = scan.run(scan.tune.__class__(block=b, chunk=c, grid=0, refill=r, use_graph=0, fused=f))
        np.testing.assert_array_equal(t, ref)
    t, _, _ = scan.run(best)
    np.testing.assert_array_equal(t, ref)


def test_two_phase_and_single_kernel_work_counters_agree(ctx):
    """Both batch paths walk the same faces and take the same RK4 steps."""
    bls, n, seeds, scan = small_scan(ctx)
    T = scan.tune.__class__
    t2, w2, _ = scan.run(T(block=0, chunk=128, grid=0, refill=16, use_graph=0, fused=0))
    t1, w1, _ = scan.run(T(block=256, chunk=128, grid=0, refill=16, use_graph=0, fused=1))
    np.testing.assert_array_equal(t1, t2)
    np.testing.assert_array_equal(w1[:2], w2[:2])
    assert w1[1] > 0


def test_config2_scan_runs_and_matches_sampled_configurations(ctx):
    """The 256-configuration scan; three configurations are re-run one by one."""
    bls, n, seeds = configs.config2()
    scan = BatchScan(bls, n, seeds, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, ctx=ctx)
    tally, work, ms = scan.run(graph=True)
    assert tally.sum() == n.sum()
    frac = scan.detected_fraction(tally)
    print("configs", len(bls), "molecules", int(n.sum()), "ms", ms, "detected fraction range", frac.min(), frac.max())
    for c in (0, int(np.argmin(n)), int(np.argmax(frac))):
        eng = Engine(bls[c], ctx)
        x, y, z, vx, vy, vz = [a.to_host() for a in eng.draw_gaussian(int(n[c]), configs.SOURCE_MEAN,
                                                                          configs.SOURCE_SIGMA, int(seeds[c]))]
        r = eng.propagate(x, y, z, vx, vy, vz, outputs=False)
        want = dict(zip(eng.labels, r["tally"].sum(axis=0).tolist()))
        assert as_dict(scan, tally, c) == {k: v for k, v in want.items() if v}, c


def test_batch_with_an_unfolded_lens_equals_per_configuration_propagation(ctx):
    """A configuration whose lens is at V0 = 0 (force scale 0, not folded into the step constants)
    keeps the batch on the two-phase path: that lens takes the plain RK4 step for its own molecules and
    the others the folded one, in both paths; tallies equal per-configuration propagation and the
    oracle."""
    bls = [configs.config1(V0=30e3), configs.config1(V0=0.0), configs.config1(V0=18e3, r3=3e-3)]
    n = np.array([9_000, 6_000, 4_321])
    seeds = np.arange(len(bls), dtype=np.uint64) + np.uint64(777)
    scan = BatchScan(bls, n, seeds, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, ctx=ctx)
    assert scan.split_path
    T = scan.tune.__class__
    dev, orc = per_config_reference(ctx, scan, bls, n, seeds)
    for fused in (0, 1):
        tally, work, ms = scan.run(T(block=0, chunk=128, grid=0, refill=0, use_graph=0, fused=fused))
        assert tally.sum() == n.sum()
        for c in range(len(bls)):
            got = as_dict(scan, tally, c)
            assert got == {k: v for k, v in dev[c].items() if v}, (fused, c, got, dev[c])
            assert got == {k: v for k, v in orc[c].items() if v}, (fused, c)


def test_batch_beyond_one_slice_equals_per_configuration_propagation(ctx):
    """A batch of more than 2^27 molecules runs the two-phase path in slices that reuse the queue
    (a configuration straddles the slice boundary): per-configuration tallies equal ctraj_propagate's
    over the same device draws, and the work counters equal the single kernel's."""
    bls = [configs.config1(V0=30e3), configs.config1(V0=22e3, lens_z0=0.75, r3=3e-3)]
    n = np.array([70_000_000, 70_000_001])
    seeds = np.arange(len(bls), dtype=np.uint64) + np.uint64(4242)
    scan = BatchScan(bls, n, seeds, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, ctx=ctx)
    assert scan.split_path and scan.total > (1 << 27)
    T = scan.tune.__class__
    tally, work, _ = scan.run(T(block=0, chunk=128, grid=0, refill=0, use_graph=0, fused=0))
    assert tally.sum() == n.sum()
    for c, bl in enumerate(bls):
        eng = Engine(bl, ctx)
        cols = eng.draw_gaussian(int(n[c]), configs.SOURCE_MEAN, configs.SOURCE_SIGMA, int(seeds[c]))
        inp = dict(zip(("x", "y", "z", "vx", "vy", "vz"), cols))
        tal = ctx.alloc(eng.cb.n_labels, np.int64).fill(0)
        eng.launch(inp, tally=tal, n=int(n[c]))
        want = dict(zip(eng.labels, tal.to_host().tolist()))
        del cols, inp
        assert as_dict(scan, tally, c) == {k: v for k, v in want.items() if v}, c
    t1, w1, _ = scan.run(T(block=256, chunk=128, grid=0, refill=16, use_graph=0, fused=1))
    np.testing.assert_array_equal(t1, tally)
    np.testing.assert_array_equal(w1[:4], work[:4])


def random_configurations(seed: int, n_cfg: int = 24):
    """Configurations with rectangular apertures and field plates besides circular ones, lenses of
    random voltage, position, bore, length and dt (hexapole spline, RK4), one of them without a lens."""
    from trajectories.beamline_elements import ElectrostaticLens, FieldPlates, RectangularAperture

    rs = np.random.RandomState(seed)
    stark = configs.stark_table()
    bls = []
    for c in range(n_cfg):
        z, els = 0.05, []

        def aperture(name):
            nonlocal z
            kind = rs.randint(3)
            L = 0.0 if rs.rand() < 0.5 else rs.uniform(1e-3, 0.05)
            if kind == 0:
                el = CircularAperture(name, z, L, d=rs.uniform(4e-3, 30e-3))
            elif kind == 1:
                el = RectangularAperture(name, z, L, x0=rs.uniform(-1e-3, 1e-3), y0=rs.uniform(-1e-3, 1e-3),
                                         w=rs.uniform(4e-3, 30e-3), h=rs.uniform(4e-3, 30e-3))
            else:
                el = FieldPlates(name, z, rs.uniform(0.05, 0.6), x0=rs.uniform(-1e-3, 1e-3), w=rs.uniform(6e-3, 30e-3))
            z = el.z0 + el.L + rs.uniform(0.01, 0.4)
            return el

        for k in range(rs.randint(0, 3)):
            els.append(aperture(f"before {k}"))
        if c != 3:
            lens = ElectrostaticLens("ES lens", z, rs.uniform(0.2, 0.8), d=rs.uniform(0.02, 0.05),
                                     V=rs.uniform(5e3, 40e3), stark=stark, field="hexapole", integrator="rk4",
                                     dt=10e-6)
            els.append(lens)
            z = lens.z1 + rs.uniform(0.05, 0.4)
        for k in range(rs.randint(1, 4)):
            els.append(aperture(f"after {k}"))
        bls.append(Beamline(els, configs.G))
    n = rs.randint(0, 30_000, n_cfg)
    n[5] = 0
    seeds = rs.randint(0, 2 ** 31, n_cfg).astype(np.uint64)
    return bls, n, seeds


@pytest.mark.parametrize("fused", [False, True])
def test_random_configurations_with_all_element_kinds_equal_per_configuration_runs(ctx, fused):
    """The general (all element kinds) instantiations of the batched kernels, split and fused:
    per-configuration tallies equal ctraj_propagate and the oracle on the same device draws."""
    bls, n, seeds = random_configurations(77)
    scan = BatchScan(bls, n, seeds, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, ctx=ctx)
    T = scan.tune.__class__
    tally, work, ms = scan.run(T(block=256 if fused else 0, chunk=128, grid=0, refill=8 if not fused else 16,
                                 use_graph=0, fused=int(fused)))
    assert tally.sum() == n.sum()
    dev, orc = per_config_reference(ctx, scan, bls, n, seeds)
    for c in range(len(bls)):
        got = as_dict(scan, tally, c)
        assert got == {k: v for k, v in dev[c].items() if v}, (c, got, dev[c])
        assert got == {k: v for k, v in orc[c].items() if v}, (c, got, orc[c])
