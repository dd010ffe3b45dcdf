"""GPU: the benchmark workloads at their full sizes, through properties that hold at any size.

The CPU oracle is too slow for 10^7-10^8 molecules inside a test, so at full size the checks are:
every molecule gets exactly one label (tallies sum to n); the two-phase path and the single-kernel
lane walker give the same bits for every molecule; random samples of molecules equal the oracle bit
for bit; kept-molecule indices are exactly the detected molecules; per-state tallies sum to the
state draw counts; recorded z-grid samples equal the oracle's resampling of its own rows.
"""
import numpy as np
import pytest

from oracle import oracle as O
from trajectories import _native as N
from trajectories import configs
from trajectories.engine import Engine, draw_gaussian_at, draw_uniform_u8
from trajectories.stream import StreamedRun

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "t")


def sample_equals_oracle(eng, X, V, dev, idx):
    o = O.propagate(eng.cb, X[0][idx], X[1][idx], X[2][idx], V[0][idx], V[1][idx], V[2][idx])
    np.testing.assert_array_equal(dev["label"][idx], o["label"])
    for k in FIELDS:
        a, b = dev["out"][k][idx], o["out"][k]
        assert ((a == b) | (np.isnan(a) & np.isnan(b))).all(), k


def test_config1_at_1e6_every_molecule_equals_the_oracle(ctx):
    """The ES-lens beamline at its parity size, 10^6 molecules: every molecule's label, status and final
    phase space, the tallies and the work counters equal the (multithreaded) oracle bit for bit."""
    n = 1_000_000
    X, V = configs.draw_source(n, seed=1)
    eng = Engine(configs.config1(), ctx)
    dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], want_accel=True, aoi_labels=("Detected",))
    orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2])
    np.testing.assert_array_equal(dev["label"], orc["label"])
    np.testing.assert_array_equal(dev["status"], orc["status"])
    for k in FIELDS:
        a, b = dev["out"][k], orc["out"][k]
        assert ((a == b) | (np.isnan(a) & np.isnan(b))).all(), k
    for q in range(3):
        np.testing.assert_array_equal(dev["out_a"][q], orc["out_a"][q])
    np.testing.assert_array_equal(dev["tally"][0], np.bincount(orc["label"], minlength=len(eng.labels)))
    np.testing.assert_array_equal(dev["work"][:2], orc["work"][:2])
    det = np.nonzero(orc["label"] == eng.cb.label_detected)[0]
    o = np.argsort(dev["aoi_index"])
    np.testing.assert_array_equal(dev["aoi_index"][o], det)
    for k in FIELDS:
        np.testing.assert_array_equal(dev["aoi_out"][k][o], orc["out"][k][det])


def test_config1_at_1e7_two_phase_equals_lane_walker_and_samples_equal_oracle(ctx):
    n = 10_000_000
    X, V = configs.draw_source(n, seed=1)
    eng = Engine(configs.config1(), ctx)
    dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], aoi_labels=("Detected",))
    assert int(dev["tally"].sum()) == n
    lanes = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], flags=N.RUN_LANES, aoi_labels=("Detected",))
    np.testing.assert_array_equal(dev["label"], lanes["label"])
    for k in FIELDS:
        np.testing.assert_array_equal(dev["out"][k], lanes["out"][k])
    np.testing.assert_array_equal(dev["tally"], lanes["tally"])
    np.testing.assert_array_equal(dev["work"][:4], lanes["work"][:4])
    assert 0 < int(dev["work"][N.WORK_FACES_POST]) < int(dev["work"][N.WORK_FACES])
    assert int(lanes["work"][N.WORK_FACES_POST]) == 0
    det = np.nonzero(dev["label"] == eng.cb.label_detected)[0]
    assert det.size > 10_000
    np.testing.assert_array_equal(np.sort(dev["aoi_index"]), det)
    rng = np.random.RandomState(11)
    sample_equals_oracle(eng, X, V, dev, np.unique(rng.randint(0, n, 4000)))
    sample_equals_oracle(eng, X, V, dev, det[rng.randint(0, det.size, 300)])


def test_config0_at_1e8_tallies_kept_molecules_and_samples(ctx):
    n = 100_000_000
    X, V = configs.draw_source(n, seed=0)
    eng = Engine(configs.config0(), ctx)
    dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], aoi_labels=("Detected",))
    assert int(dev["tally"].sum()) == n
    np.testing.assert_array_equal(dev["tally"][0], np.bincount(dev["label"], minlength=eng.cb.n_labels))
    det = np.nonzero(dev["label"] == eng.cb.label_detected)[0]
    o = np.argsort(dev["aoi_index"])
    np.testing.assert_array_equal(dev["aoi_index"][o], det)
    for k in FIELDS:
        np.testing.assert_array_equal(dev["aoi_out"][k][o], dev["out"][k][det])
    rng = np.random.RandomState(12)
    sample_equals_oracle(eng, X, V, dev, np.unique(rng.randint(0, n, 200_000)))


def test_config4_streamed_at_1e8_state_tallies_detected_set_and_samples(ctx):
    n, seed = 100_000_000, 4
    run = StreamedRun(configs.config4(), n, seed=seed, ctx=ctx)
    res = run.run(trajectories=True)
    det_label = res.labels.index("Detected")
    assert int(res.tally.sum()) == n
    st = draw_uniform_u8(ctx, n, 4, seed).to_host()
    np.testing.assert_array_equal(res.tally.sum(axis=1), np.bincount(st, minlength=4))
    d = res.detected_index
    assert d.size == int(res.tally[:, det_label].sum()) > 100_000
    assert d[0] >= 0 and d[-1] < n and np.all(np.diff(d) > 0)
    np.testing.assert_array_equal(res.detected_state, st[d])
    assert np.all(res.samples_written > 0)
    # a sample of detected molecules, regenerated from the same counter-based streams, through the oracle
    rng = np.random.RandomState(13)
    js = np.unique(rng.randint(0, d.size, 24))
    sel = d[js]
    d_idx = ctx.to_device(np.ascontiguousarray(sel))
    cols = [a.to_host()[: sel.size] for a in draw_gaussian_at(ctx, d_idx, sel.size, configs.SOURCE_MEAN,
                                                              configs.SOURCE_SIGMA, seed)]
    eng = run.engine
    ro = O.propagate(eng.cb, *cols, state=st[sel], rows_cap=1400)
    np.testing.assert_array_equal(ro["label"], np.full(sel.size, eng.cb.label_detected))
    got = res.trajectories_of(js)
    for q in range(sel.size):
        want = O.zgrid_from_rows(ro["rows"][q], int(ro["n_rows"][q]), run.z_start, run.z_step, res.n_samples).T
        same = (got[q] == want) | (np.isnan(got[q]) & np.isnan(want))
        assert same.all(), (int(sel[q]), np.argwhere(~same)[:5])


def test_a_run_of_several_slices_equals_the_lane_walker(ctx):
    """More than 2^26 molecules: the split kernels run the lens-entry queue slice by slice inside one
    ctraj_propagate; every molecule's label and final phase space, the tallies, the work counters and the
    kept molecules equal the single-kernel lane walker's (which needs no queue)."""
    from trajectories.engine import draw_gaussian

    n = (1 << 26) + 1_000_003
    eng = Engine(configs.config1(), ctx)
    xs = draw_gaussian(ctx, n, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, seed=123)
    inp = dict(zip(("x", "y", "z", "vx", "vy", "vz"), xs))
    res = []
    for flags in (0, N.RUN_LANES):
        tally = ctx.alloc(eng.cb.n_labels, np.int64).fill(0)
        work = ctx.alloc(N.WORK_N, np.uint64).fill(0)
        label = ctx.alloc(n, np.int32)
        out = {k: ctx.alloc(n) for k in ("x", "y", "t")}
        count = ctx.alloc(1, np.int64).fill(0)
        cap = n // 50
        index = ctx.alloc(cap, np.int64)
        kept = {k: ctx.alloc(cap) for k in FIELDS}
        eng.launch(inp, out=out, label=label, tally=tally, work=work, aoi_mask=1 << eng.cb.label_detected,
                   aoi_cap=cap, aoi_count=count, aoi_index=index, aoi_out=kept, n=n, flags=flags)
        ctx.sync()
        cnt = int(count.to_host()[0])
        assert 0 < cnt <= cap
        o = np.argsort(index.to_host()[:cnt])
        res.append({"tally": tally.to_host(), "work": work.to_host(), "label": label.to_host(),
                    "out": {k: v.to_host() for k, v in out.items()}, "index": index.to_host()[:cnt][o],
                    "kept": {k: v.to_host()[:cnt][o] for k, v in kept.items()}})
        del label, out, kept
    a, b = res
    assert int(a["tally"].sum()) == n
    np.testing.assert_array_equal(a["tally"], b["tally"])
    np.testing.assert_array_equal(a["work"][:4], b["work"][:4])
    np.testing.assert_array_equal(a["label"], b["label"])
    for k in a["out"]:
        np.testing.assert_array_equal(a["out"][k], b["out"][k])
    np.testing.assert_array_equal(a["index"], b["index"])
    for k in FIELDS:
        np.testing.assert_array_equal(a["kept"][k], b["kept"][k])


def test_a_billion_molecules_on_the_split_path_equal_the_lane_walker(ctx):
    """10^9 molecules (the reference's large SPA runs, examples/SPA/SPA_distributions.py:99-102) drawn
    on the device run the split kernels in sixteen slices inside one ctraj_propagate: the tallies, the
    work counters and every detected molecule's index and final phase space equal the single-kernel
    lane walker's."""
    from trajectories.engine import draw_gaussian

    n = 1_000_000_000
    eng = Engine(configs.config1(), ctx)
    xs = draw_gaussian(ctx, n, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, seed=2024)
    inp = dict(zip(("x", "y", "z", "vx", "vy", "vz"), xs))
    cap = n // 200
    res = []
    for flags in (0, N.RUN_LANES):
        tally = ctx.alloc(eng.cb.n_labels, np.int64).fill(0)
        work = ctx.alloc(N.WORK_N, np.uint64).fill(0)
        count = ctx.alloc(1, np.int64).fill(0)
        index = ctx.alloc(cap, np.int64)
        kept = {k: ctx.alloc(cap) for k in FIELDS}
        eng.launch(inp, tally=tally, work=work, aoi_mask=1 << eng.cb.label_detected, aoi_cap=cap, aoi_count=count,
                   aoi_index=index, aoi_out=kept, n=n, flags=flags)
        ctx.sync()
        cnt = int(count.to_host()[0])
        assert 0 < cnt <= cap
        o = np.argsort(index.to_host()[:cnt])
        res.append({"tally": tally.to_host(), "work": work.to_host(), "index": index.to_host()[:cnt][o],
                    "kept": {k: v.to_host()[:cnt][o] for k, v in kept.items()}})
        del kept, index
    a, b = res
    assert int(a["tally"].sum()) == n
    np.testing.assert_array_equal(a["tally"], b["tally"])
    np.testing.assert_array_equal(a["work"][:4], b["work"][:4])
    assert a["work"][N.WORK_FACES_POST] > 0 and b["work"][N.WORK_FACES_POST] == 0  # split kernels, lane walker
    np.testing.assert_array_equal(a["index"], b["index"])
    for k in FIELDS:
        np.testing.assert_array_equal(a["kept"][k], b["kept"][k])
