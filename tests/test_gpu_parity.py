"""GPU parity: libctraj.so (HIP, gfx950) against the CPU oracle on the same inputs.

The device and the oracle perform the same IEEE operations in the same order (contraction off,
explicit FMAs mirrored), so final phase space, labels, status bits, tallies and work counters are
required to agree bit for bit.  The stated tolerance of the workload (1e-9 relative on x and v,
exact per-element counts except molecules within 1e-9 m of an edge) is the documented fallback
and is what `assert_parity` enforces if a bit differs.
"""
import numpy as np
import pytest

from oracle import oracle as O
from trajectories import configs
from trajectories import _native as N
from trajectories.beamline import Beamline
from trajectories.beamline_elements import (CircularAperture, ElectrostaticLens, FieldPlates,
                                            RectangularAperture)
from trajectories.compile import compile_elements
from trajectories.distributions import CeNTREXPositionDistribution, CeNTREXVelocityDistribution
from trajectories.engine import Engine
from trajectories.stark import TwoLevelStark

from beamlines import reference_example_beamline

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "t")


def assert_parity(dev, orc, tag, rel=1e-9, edge_band=1e-9):
    lab_d, lab_o = dev["label"], orc["label"]
    mism = np.nonzero(lab_d != lab_o)[0]
    if mism.size:
        near = orc["edge"][mism] < edge_band
        assert near.all(), f"{tag}: {np.count_nonzero(~near)} label mismatches away from edges: {mism[~near][:10]}"
        print(f"{tag}: {mism.size} label mismatches, all within {edge_band} m of an edge: {mism[:20]}")
    ok = lab_d == lab_o
    bitwise = True
    for k in FIELDS:
        a, b = dev["out"][k][ok], orc["out"][k][ok]
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        if not same.all():
            bitwise = False
            scale = np.maximum(np.abs(b), 1e-12)
            err = np.nanmax(np.abs(a - b) / scale)
            assert err <= rel, f"{tag}: field {k} max rel diff {err:.3e} > {rel}"
    print(f"{tag}: bitwise={bitwise} n={lab_d.size} label mismatches={mism.size}")
    return bitwise


def run_both(bl, X, V, state=None, n_states=None, rows=False):
    eng = Engine(bl, n_states=n_states)
    dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], state=state, want_accel=True)
    orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2], state=state)
    return eng, dev, orc


def test_config0_ballistic_bitwise():
    bl = configs.config0()
    X, V = configs.draw_source(100_000, seed=0)
    eng, dev, orc = run_both(bl, X, V)
    assert assert_parity(dev, orc, "cfg0 1e5"), "ballistic walker must be bit-exact"
    cnt = np.bincount(orc["label"], minlength=len(eng.labels))
    np.testing.assert_array_equal(dev["tally"][0], cnt)
    assert dev["tally"].sum() == 100_000
    assert int(dev["work"][N.WORK_FACES]) == int(orc["work"][0])
    for q in range(3):
        np.testing.assert_array_equal(dev["out_a"][q], orc["out_a"][q])


def test_config1_lens_rk4_bitwise():
    bl = configs.config1()
    X, V = configs.draw_source(200_000, seed=1)
    eng, dev, orc = run_both(bl, X, V)
    bit = assert_parity(dev, orc, "cfg1 2e5")
    np.testing.assert_array_equal(dev["tally"][0], np.bincount(orc["label"], minlength=len(eng.labels)))
    assert int(dev["work"][N.WORK_STEPS]) == int(orc["work"][1])
    np.testing.assert_array_equal(dev["status"], orc["status"])
    assert bit


def test_config1_focusing_sign():
    bl = configs.config1(stark=configs.stark_table(+1.0))
    X, V = configs.draw_source(100_000, seed=11)
    eng, dev, orc = run_both(bl, X, V)
    assert assert_parity(dev, orc, "cfg1 focusing")
    det = eng.labels.index("Detected")
    assert dev["tally"][0][det] > 0.01 * 100_000  # the + sign focuses: far more reach the detector


def test_reference_mode_example_beamline():
    bl = reference_example_beamline()
    rs = np.random.RandomState(5)
    n = 50_000
    V = np.array([rs.normal(0, 39.5, n), rs.normal(0, 39.5, n), rs.normal(184, 16, n)])
    th = rs.uniform(size=n) * 2 * np.pi
    r = np.sqrt(rs.uniform(size=n)) * 0.01
    X = np.array([r * np.cos(th), r * np.sin(th), np.full(n, 0.25 * 0.0254)])
    # bias a third of the molecules onto the axis so that many reach the lens and the plates
    X[:2, : n // 3] *= 0.05
    V[:2, : n // 3] *= 0.02
    eng, dev, orc = run_both(bl, X, V)
    assert assert_parity(dev, orc, "ref example")
    np.testing.assert_array_equal(dev["status"], orc["status"])
    labs = set(np.array(eng.labels)[np.unique(dev["label"])])
    assert {"ES lens", "Lens entrance"} & labs or "Inside lens" in labs
    assert "Detected" in labs or "Field plates" in labs


def test_multistate_config4_tables():
    bl = configs.config4()
    X, V = configs.draw_source(100_000, seed=4)
    st = np.random.RandomState(44).randint(0, 4, 100_000).astype(np.uint8)
    eng, dev, orc = run_both(bl, X, V, state=st, n_states=4)
    assert assert_parity(dev, orc, "cfg4 4 states")
    for s in range(4):
        m = st == s
        np.testing.assert_array_equal(dev["tally"][s], np.bincount(orc["label"][m], minlength=len(eng.labels)))


def test_record_rows_match_oracle():
    for bl, seed in ((configs.config1(), 21), (reference_example_beamline(), 22), (configs.config0(), 23)):
        X, V = configs.draw_source(2000, seed=seed)
        if bl.elements[0].name == "4K shield":
            X[:2] *= 0.05
            V[:2] *= 0.05
        eng = Engine(bl)
        rows, nrows, lab = eng.record_rows(X[0], X[1], X[2], V[0], V[1], V[2], cap=96)
        cap = rows.shape[1]
        orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2], rows_cap=cap)
        np.testing.assert_array_equal(nrows, orc["n_rows"])
        np.testing.assert_array_equal(lab, orc["label"])
        np.testing.assert_array_equal(rows, orc["rows"])


def test_zgrid_samples_match_post_processing_rule():
    bl = configs.config1(stark=configs.stark_table(+1.0))
    X, V = configs.draw_source(1000, seed=31)
    eng = Engine(bl)
    K = 2501
    samp, written, lab = eng.zgrid(X[0], X[1], X[2], V[0], V[1], V[2], z_start=0.0, z_step=1e-3, n_samples=K)
    orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2], rows_cap=512)
    for i in range(0, 1000, 7):
        ref = O.zgrid_from_rows(orc["rows"][i], orc["n_rows"][i], 0.0, 1e-3, K)
        np.testing.assert_array_equal(samp[:, :, i], ref)
        assert written[i] == np.count_nonzero(~np.isnan(ref[0]))


def test_philox_source_matches_oracle(ctx):
    from trajectories.engine import draw_gaussian

    n = 100_003
    mean = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 184.0])
    sig = np.array([1.5e-3, 1.5e-3, 0.0, 1.5, 1.5, 16.0])
    dev = [d.to_host() for d in draw_gaussian(ctx, n, mean, sig, seed=123456789, offset=77)]
    ref = O.draw_gaussian(n, mean, sig, seed=123456789, offset=77)
    for d in range(6):
        # uniforms are bit-identical; log/cos/sin of the two math libraries may differ in the last bits
        np.testing.assert_allclose(dev[d], ref[d], rtol=1e-13, atol=1e-13 * max(sig[d], 1e-300))
    z = (dev[5] - 184.0) / 16.0
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02


def test_aoi_compaction_and_first_index():
    bl = configs.config0()
    X, V = configs.draw_source(100_000, seed=7)
    eng = Engine(bl)
    res = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], aoi_labels=("Detected", "Aperture 3"))
    want = np.nonzero(np.isin(res["label"], [eng.labels.index("Detected"), eng.labels.index("Aperture 3")]))[0]
    np.testing.assert_array_equal(np.sort(res["aoi_index"]), want)
    o = np.argsort(res["aoi_index"])
    np.testing.assert_array_equal(res["aoi_label"][o], res["label"][want])
    for k in FIELDS:
        np.testing.assert_array_equal(res["aoi_out"][k][o], res["out"][k][want])
    for i, name in enumerate(eng.labels):
        hit = np.nonzero(res["label"] == i)[0]
        if hit.size:
            assert res["first_index"][i] == hit[0]


def test_run_simulation_dropin_counts():
    from trajectories.trajectory_simulator import TrajectorySimulator

    bl = configs.config0()
    xdist, vdist = configs.source()
    np.random.seed(3)
    sim = TrajectorySimulator()
    sim.run_simulation(bl, "r", vdist=vdist, xdist=xdist, N_traj=20_000, apertures_of_interest=["Detected"], n_jobs=2)
    X, V = configs.draw_source(20_000, seed=3, n_jobs=2)
    orc = O.propagate(compile_elements(bl.elements, bl.g), X[0], X[1], X[2], V[0], V[1], V[2])
    labels = compile_elements(bl.elements, bl.g).labels
    want = {}
    for lab in orc["label"]:  # first-seen insertion order, as Counter.increment_counter builds it
        want[labels[lab]] = want.get(labels[lab], 0) + 1
    assert sim.counter.counter_dict == want
    assert list(sim.counter.counter_dict) == list(want)
    det = np.nonzero(orc["label"] == labels.index("Detected"))[0]
    assert len(sim.result.molecules) == det.size
    m = sim.result.molecules[3]
    assert m.aperture_hit == "Detected" and m.alive
    np.testing.assert_array_equal(m.x()[0], orc["out"]["x"][det[3]])


def test_beamline_propagate_single_molecule_smoke_kat():
    """apertures.py:289-306: a d=0 aperture at z=0.99 stops the default molecule (gravity moves it off axis)."""
    from trajectories.molecule import Molecule

    ap = CircularAperture(name="40K Shield", z0=0.99, L=0.01, d=0)
    rect = RectangularAperture(name="Detection aperture", z0=100, L=0.01, w=0.00001, h=1e-6)
    bl = Beamline([ap, rect])
    m = Molecule()
    m.init_trajectory(bl)
    bl.propagate_through(m)
    assert not m.alive and m.aperture_hit == "40K Shield"
    assert m.trajectory.x.shape[0] == 2
    dt = 0.99 / 200
    assert m.trajectory.x[1, 1] == pytest.approx(-9.80665 * dt * dt / 2, rel=1e-15)


def test_lane_refill_kernel_matches_simple_kernel():
    """The split kernels (in either claim order) and the single-kernel lane-refill walker only reorder
    work: every refill threshold, all of them and the one-thread-per-molecule kernel give identical bits, in
    configuration mode (one and four lens states) and in reference mode.  Molecules stopped inside
    the lens are kept too, so the lens phase's own output path is compared."""
    cases = ((configs.config1(), 41, None), (reference_example_beamline(), 42, None), (configs.config4(), 43, 4))
    for bl, seed, ns in cases:
        X, V = configs.draw_source(50_000, seed=seed)
        if bl.elements[0].name == "4K shield":
            X[:2] *= 0.05
            V[:2] *= 0.05
        st = None if ns is None else np.random.RandomState(seed).randint(0, ns, X.shape[1]).astype(np.uint8)
        eng = Engine(bl, n_states=ns)
        keep = ("Detected", "Inside lens")
        base = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], state=st, flags=N.RUN_SIMPLE, want_accel=True,
                             aoi_labels=keep)
        post_faces = set()
        ref_mode = eng.cb.integrator == N.INTEG_REF
        # flags 0: the lens kernel takes hexapole lens molecules longest-lived first; RUN_QUEUE_ORDER in
        # the order the walk queued them
        runs = ((0, 1, True), (0, 24, True), (0, 64, True), (0, 1, False), (0, 12, False), (0, 64, False),
                (N.RUN_QUEUE_ORDER, 1, True), (N.RUN_QUEUE_ORDER, 8, False),
                (N.RUN_LANES, 1, True), (N.RUN_LANES, 16, True), (N.RUN_LANES, 64, True))
        for flags, refill, accel in runs:
            res = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], state=st, refill=refill, want_accel=accel,
                                aoi_labels=keep, flags=flags)
            np.testing.assert_array_equal(res["label"], base["label"])
            np.testing.assert_array_equal(res["status"], base["status"])
            for k in FIELDS:
                np.testing.assert_array_equal(res["out"][k], base["out"][k])
            if accel:
                for q in range(3):
                    np.testing.assert_array_equal(res["out_a"][q], base["out_a"][q])
            np.testing.assert_array_equal(res["tally"], base["tally"])
            np.testing.assert_array_equal(res["first_index"], base["first_index"])
            np.testing.assert_array_equal(res["work"][:4], base["work"][:4])
            # the split kernels count the faces after the lens apart, the same at every threshold; a
            # reference-mode lens asked for stored accelerations runs the lane walker
            if flags in (0, N.RUN_QUEUE_ORDER) and not (ref_mode and accel):
                post_faces.add(int(res["work"][N.WORK_FACES_POST]))
            else:
                assert int(res["work"][N.WORK_FACES_POST]) == 0
            o = np.argsort(res["aoi_index"])
            ob = np.argsort(base["aoi_index"])
            np.testing.assert_array_equal(res["aoi_index"][o], base["aoi_index"][ob])
            np.testing.assert_array_equal(res["aoi_label"][o], base["aoi_label"][ob])
            for k in FIELDS:
                np.testing.assert_array_equal(res["aoi_out"][k][o], base["aoi_out"][k][ob])
        assert len(post_faces) == 1 and 0 < post_faces.pop() < int(base["work"][N.WORK_FACES])


def test_small_and_empty_runs():
    bl = configs.config1()
    eng = Engine(bl)
    for n in (0, 1, 63, 65, 129):
        X, V = configs.draw_source(200, seed=50 + n)
        X, V = X[:, :n], V[:, :n]
        dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2])
        assert dev["tally"].sum() == n
        if n:
            orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2])
            np.testing.assert_array_equal(dev["label"], orc["label"])
            for k in FIELDS:
                np.testing.assert_array_equal(dev["out"][k], orc["out"][k])


def test_aperture_edge_decisions_bitwise():
    """Molecules that reach circular apertures at radii within a few ulps of d/2 (and at d/2 exactly,
    along x, y and the diagonal): every stop/pass decision equals the oracle's sqrt(x^2 + y^2) > d/2,
    with the device's squared-radius thresholds.  Radii chosen so that x^2 + y^2 rounds onto both
    sides of the threshold."""
    halves = [3e-3, 4e-3, 5e-3, 6e-3, 1e-3 / 3, 0.0123456789, 2.5e-3 * (1 + 2 ** -40)]
    xs, ys, ds = [], [], []
    for d in halves:
        for k in range(-6, 7):
            r = d
            for _ in range(abs(k)):
                r = np.nextafter(r, np.inf if k > 0 else -np.inf)
            xs += [r, 0.0, r / np.sqrt(2.0), -r]
            ys += [0.0, r, r / np.sqrt(2.0), 0.0]
            ds += [d] * 4
    xs, ys, ds = np.array(xs), np.array(ys), np.array(ds)
    n_per = xs.size // len(halves)
    for j, d in enumerate(halves):
        sl = slice(j * n_per, (j + 1) * n_per)
        bl = Beamline([CircularAperture("Ap", 0.5, 0.0, d=2 * d)], configs.G)
        n = n_per
        # start on the aperture plane with no transverse motion, so the tested radius is the start's
        X = np.array([xs[sl], ys[sl], np.full(n, 0.5)])
        V = np.array([np.zeros(n), np.zeros(n), np.full(n, 184.0)])
        eng, dev, orc = run_both(bl, X, V)
        np.testing.assert_array_equal(dev["label"], orc["label"])
        beyond = np.sqrt(xs[sl] ** 2 + ys[sl] ** 2) > d
        assert beyond.any() and (~beyond).any()
        det = eng.cb.label_detected
        np.testing.assert_array_equal(orc["label"] != det, beyond)


def test_aoi_capacity_smaller_than_the_kept_molecules(ctx):
    """Kept molecules beyond aoi_cap are counted but not stored; the stored ones are distinct kept
    molecules with their own phase space.  A small capacity also overflows the per-XCD staging
    parts, so both the parts and the direct path to the output are exercised; a second launch
    appends after the first."""
    bl = configs.config0()
    X, V = configs.draw_source(200_000, seed=11)
    eng = Engine(bl, ctx)
    ref = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], aoi_labels=("Detected",))
    det = eng.labels.index("Detected")
    want = np.nonzero(ref["label"] == det)[0]
    assert want.size > 600
    inp = {k: ctx.to_device(np.asarray(a)) for k, a in zip(FIELDS[:6], (X[0], X[1], X[2], V[0], V[1], V[2]))}
    for cap in (150, want.size - 1, want.size, 2 * want.size + 10):
        count = ctx.alloc(1, np.int64).fill(0)
        index = ctx.alloc(cap, np.int64).fill(-1)
        out = {k: ctx.alloc(cap) for k in FIELDS}
        for rep in range(2):  # the second launch appends
            eng.launch(inp, aoi_mask=1 << det, aoi_cap=cap, aoi_count=count, aoi_index=index, aoi_out=out,
                       n=X.shape[1])
        ctx.sync()
        assert int(count.to_host()[0]) == 2 * want.size
        got = index.to_host()
        stored = min(cap, 2 * want.size)
        assert np.all(got[:stored] >= 0)
        vals, cnt = np.unique(got[:stored], return_counts=True)
        assert np.all(np.isin(vals, want)) and np.all(cnt <= 2)
        for k in FIELDS:
            np.testing.assert_array_equal(out[k].to_host()[:stored], ref["out"][k][got[:stored]])


def test_unfolded_lens_at_zero_voltage_bitwise():
    """A lens at V0 = 0 has force scale ks = 0, which cannot be folded into the step constants
    (gravity would enter as -g/ks): it takes the lane walker with the plain RK4 step on every path
    (default, lanes, one molecule per thread), bit-identical to the oracle, and its molecules fall
    under gravity alone."""
    bl = configs.config1(V0=0.0)
    X, V = configs.draw_source(30_000, seed=44)
    eng = Engine(bl)
    assert eng.cb.elems[1].p[4] == 0.0
    orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2])
    for flags in (0, N.RUN_LANES, N.RUN_SIMPLE):
        dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], want_accel=True, flags=flags)
        assert assert_parity(dev, orc, f"V0=0 flags={flags}"), "unfolded lens step must be bit-exact"
        np.testing.assert_array_equal(dev["tally"][0], np.bincount(orc["label"], minlength=len(eng.labels)))
    det = orc["label"] == eng.cb.label_detected
    assert det.any()
    # no lens force: vx is unchanged and vy fell by g t over the flight
    np.testing.assert_allclose(orc["out"]["vx"][det], V[0][det], rtol=0, atol=1e-12)
    np.testing.assert_allclose(orc["out"]["vy"][det], V[1][det] - configs.G * orc["out"]["t"][det], rtol=0, atol=1e-9)


def launch_min(eng, X, V, n):
    """Positions and velocities in HBM; tallies, work counters and kept (detected) molecules only."""
    ctx = eng.ctx
    inp = {k: ctx.to_device(a) for k, a in zip(("x", "y", "z", "vx", "vy", "vz"), (X[0], X[1], X[2], V[0], V[1], V[2]))}
    tally = ctx.alloc(eng.cb.n_labels, np.int64).fill(0)
    work = ctx.alloc(N.WORK_N, np.uint64).fill(0)
    count = ctx.alloc(1, np.int64).fill(0)
    index = ctx.alloc(max(n, 1), np.int64)
    out = {k: ctx.alloc(max(n, 1)) for k in FIELDS}
    eng.launch(inp, tally=tally, work=work, aoi_mask=1 << eng.cb.label_detected, aoi_cap=n, aoi_count=count,
               aoi_index=index, aoi_out=out, n=n)
    ctx.sync()
    cnt = int(count.to_host()[0])
    return {"tally": tally.to_host().reshape(1, -1), "work": work.to_host(), "aoi_index": index.to_host()[:cnt],
            "aoi_out": {k: v.to_host()[:cnt] for k, v in out.items()}}


@pytest.mark.parametrize("n", [1, 777, 300_001, 2_000_003])
def test_positions_and_velocities_only_runs_equal_the_general_walker(n):
    """A run that passes only positions and velocities and asks for no per-molecule outputs takes
    the walkers that prefetch two molecules ahead (the ballistic walk, and the lens beamline's first
    phase); its tallies, work counters and kept molecules equal the general walker's (a run with
    per-molecule outputs) and the oracle's, at sizes that end the two-deep loop on either buffer."""
    for bl, seed in ((configs.config0(), 51), (configs.config1(), 52)):
        X, V = configs.draw_source(-(-n // 100) * 100, seed=seed)  # the source draws in hundreds
        X, V = np.ascontiguousarray(X[:, :n]), np.ascontiguousarray(V[:, :n])
        eng = Engine(bl)
        fast = launch_min(eng, X, V, n)
        gen = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], aoi_labels=("Detected",))
        np.testing.assert_array_equal(fast["tally"], gen["tally"])
        np.testing.assert_array_equal(fast["work"], gen["work"])
        o, og = np.argsort(fast["aoi_index"]), np.argsort(gen["aoi_index"])
        np.testing.assert_array_equal(fast["aoi_index"][o], gen["aoi_index"][og])
        for k in FIELDS:
            np.testing.assert_array_equal(fast["aoi_out"][k][o], gen["aoi_out"][k][og])
        if n <= 300_001:
            orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2])
            np.testing.assert_array_equal(fast["tally"][0], np.bincount(orc["label"], minlength=len(eng.labels)))
            det = np.nonzero(orc["label"] == eng.cb.label_detected)[0]
            np.testing.assert_array_equal(fast["aoi_index"][o], det)
            for k in FIELDS:
                np.testing.assert_array_equal(fast["aoi_out"][k][o], orc["out"][k][det])
