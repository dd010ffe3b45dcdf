"""GPU: streamed runs (device source, four lens states, chunked) and their saved z-grid trajectories.

A streamed run over several chunks must give the per-state tallies and the detected molecules of
one propagation over the same molecules (device draws copied to the host and run through the CPU
oracle), and the recorded samples must equal the oracle's post-processing resampling of its own
trajectory rows, bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle as O
from trajectories import _native as N
from trajectories import configs
from trajectories.engine import (Engine, draw_gaussian, draw_gaussian_at, draw_uniform_u8, draw_uniform_u8_at, gather,
                                 sort_i64)
from trajectories.stream import StreamedRun

pytestmark = pytest.mark.gpu


def test_state_draw_matches_oracle_and_is_uniform(ctx):
    n = 200_003
    d = draw_uniform_u8(ctx, n, 4, seed=99, offset=12345).to_host()
    o = O.draw_uniform_u8(n, 4, 99, 12345)
    np.testing.assert_array_equal(d, o)
    counts = np.bincount(d, minlength=4)
    assert counts.sum() == n and np.all(np.abs(counts - n / 4) < 5 * np.sqrt(n / 4))
    d3 = draw_uniform_u8(ctx, 1000, 3, seed=5).to_host()
    assert d3.max() == 2


def test_index_list_draws_equal_the_contiguous_streams(ctx):
    n = 5000
    full = [a.to_host() for a in draw_gaussian(ctx, n, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, 77)]
    fst = draw_uniform_u8(ctx, n, 4, 77).to_host()
    idx = np.array([4999, 0, 17, 17, 2500, 1234], np.int64)
    d_idx = ctx.to_device(idx)
    part = [a.to_host()[: idx.size] for a in draw_gaussian_at(ctx, d_idx, idx.size, configs.SOURCE_MEAN,
                                                               configs.SOURCE_SIGMA, 77)]
    for f in range(6):
        np.testing.assert_array_equal(part[f], full[f][idx])
    np.testing.assert_array_equal(draw_uniform_u8_at(ctx, d_idx, idx.size, 4, 77).to_host()[: idx.size], fst[idx])


@pytest.mark.parametrize("n", [1, 2, 1000, 300_001])
def test_radix_sort(ctx, n):
    rng = np.random.RandomState(n)
    keys = rng.randint(0, 1 << 40, n).astype(np.int64)
    d = ctx.to_device(keys)
    sort_i64(ctx, d, n)
    np.testing.assert_array_equal(d.to_host()[:n], np.sort(keys))


def test_gather_and_its_index_check(ctx):
    src = ctx.to_device(np.arange(100, dtype=np.float64) * 0.5)
    idx = ctx.to_device(np.array([107, 100, 199], np.int64))
    out = ctx.alloc(3)
    gather(ctx, 3, idx, 100, 100, [src], [out])
    np.testing.assert_array_equal(out.to_host(), [3.5, 0.0, 49.5])
    bad = ctx.to_device(np.array([100, 200], np.int64))
    with pytest.raises(N.NativeError):
        gather(ctx, 2, bad, 100, 100, [src], [out])


def test_streamed_run_equals_one_propagation_and_the_oracle(ctx):
    n, chunk, seed = 50_000, 16_384, 41
    bl = configs.config4()
    run = StreamedRun(bl, n, chunk=chunk, seed=seed, ctx=ctx)
    res = run.run(trajectories=True)
    # the same molecules, drawn whole
    cols = [a.to_host() for a in draw_gaussian(ctx, n, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, seed)]
    st = draw_uniform_u8(ctx, n, 4, seed).to_host()
    eng = Engine(bl, ctx, n_states=4)
    o = O.propagate(eng.cb, *cols, state=st)
    ref = np.zeros((4, eng.cb.n_labels), np.int64)
    np.add.at(ref, (st.astype(np.int64), o["label"]), 1)
    np.testing.assert_array_equal(res.tally, ref)
    det = np.nonzero(o["label"] == eng.cb.label_detected)[0]
    np.testing.assert_array_equal(res.detected_index, det)
    np.testing.assert_array_equal(res.detected_state, st[det])
    assert res.n_samples == 2501
    print("detected", res.n_detected, "per state", res.tally[:, eng.cb.label_detected], res.timings)
    # samples of some detected molecules against the oracle's resampling of its rows
    rng = np.random.RandomState(0)
    js = np.unique(rng.randint(0, res.n_detected, 40))
    got = res.trajectories_of(js)
    sel = det[js]
    ro = O.propagate(eng.cb, *[c[sel] for c in cols], state=st[sel], rows_cap=1400)
    assert np.all(ro["n_rows"] <= 1400)
    for q, j in enumerate(js):
        want = O.zgrid_from_rows(ro["rows"][q], int(ro["n_rows"][q]), 0.0, 1e-3, 2501).T
        same = (got[q] == want) | (np.isnan(got[q]) & np.isnan(want))
        assert same.all(), (j, np.argwhere(~same)[:5])
    assert np.all(res.samples_written[js] > 2400)


def test_molecule_major_zgrid_is_the_field_major_transposed(ctx):
    """CTRAJ_REC_ZGRID_MOL writes the same samples as CTRAJ_REC_ZGRID, [n][k][f] instead of [f][k][n],
    NaN entries included; a 4-state beamline with mixed states."""
    n = 3000
    cols = [a.to_host() for a in draw_gaussian(ctx, n, configs.SOURCE_MEAN, configs.SOURCE_SIGMA, 21)]
    st = draw_uniform_u8(ctx, n, 4, 21).to_host()
    eng = Engine(configs.config4(), ctx, n_states=4)
    fm, w1, l1 = eng.zgrid(*cols, state=st, n_samples=2501)
    mm, w2, l2 = eng.zgrid_by_molecule(*cols, state=st, n_samples=2501)
    np.testing.assert_array_equal(w1, w2)
    np.testing.assert_array_equal(l1, l2)
    want = fm.transpose(2, 1, 0)
    same = (mm == want) | (np.isnan(mm) & np.isnan(want))
    assert same.all()
    assert np.isfinite(mm).any() and np.isnan(mm).any()


def test_disc_source_matches_oracle(ctx):
    """ctraj_draw_disc (CeNTREXPositionDistribution on the device) against the oracle's restatement:
    the Philox words are the same, so the points agree to the last bits that the two math libraries'
    sqrt/cos/sin may differ in."""
    from trajectories.engine import draw_disc

    n, R, z0 = 100_003, 0.01, 0.00635
    dev = [a.to_host()[:n] for a in draw_disc(ctx, n, R, z0, seed=2024, offset=99)]
    ref = O.draw_disc(n, R, z0, 2024, offset=99)
    for d in range(2):
        np.testing.assert_allclose(dev[d], ref[d], rtol=1e-13, atol=1e-13 * R)
    np.testing.assert_array_equal(dev[2], ref[2])
    assert np.all(dev[0] ** 2 + dev[1] ** 2 <= R * R * (1 + 1e-15))


def test_run_simulation_on_device_draws_equals_the_host_arrays_path(ctx):
    """run_simulation(device_seed=...) draws and propagates on the device; the same molecules, drawn
    with the same keys and copied to the host, give the same Counter, kept molecules and trajectories
    through simulate_arrays, whatever the chunk size."""
    from trajectories.trajectory_simulator import TrajectorySimulator, device_stream_keys, simulate_arrays

    from trajectories.distributions import CeNTREXPositionDistribution

    bl = configs.config1()
    xdist, vdist = CeNTREXPositionDistribution(), configs.source()[1]  # the reference's default disc source
    n_traj = 250_000
    sim = TrajectorySimulator()
    keep = ["Detected", "Aperture 2"]
    sim.run_simulation(bl, "dev", vdist=vdist, xdist=xdist, N_traj=n_traj, apertures_of_interest=keep,
                       device_seed=31, chunk=1 << 16)
    res = sim.result
    n = res.run.n
    assert n == n_traj
    kx, kv = device_stream_keys(31)
    X = np.array([a.to_host()[:n] for a in xdist.draw_device(n, kx, ctx=ctx)])
    V = np.array([a.to_host()[:n] for a in vdist.draw_device(n, kv, ctx=ctx)])
    host = simulate_arrays(bl, X, V, keep, ctx=ctx)
    assert list(res.counter.counter_dict.items()) == list(host.counter.counter_dict.items())
    np.testing.assert_array_equal(res.run.tally, host.tally)
    np.testing.assert_array_equal(res.run.aoi_index, host.aoi_index)
    np.testing.assert_array_equal(res.run.aoi_label, host.aoi_label)
    assert len(res.molecules) == len(host.molecules) > 20
    np.testing.assert_array_equal(res.molecules._init, host.molecules._init)
    for j in (0, len(res.molecules) // 2, len(res.molecules) - 1):
        a, b = res.molecules[j], host.molecules[j]
        assert a.aperture_hit == b.aperture_hit
        np.testing.assert_array_equal(a.trajectory.x, b.trajectory.x)
        np.testing.assert_array_equal(a.trajectory.v, b.trajectory.v)
    # another chunking draws the same molecules
    sim2 = TrajectorySimulator()
    sim2.run_simulation(bl, "dev2", vdist=vdist, xdist=xdist, N_traj=n_traj, apertures_of_interest=keep,
                        device_seed=31, chunk=100_000)
    np.testing.assert_array_equal(sim2.result.run.tally, res.run.tally)
    np.testing.assert_array_equal(sim2.result.run.aoi_index, res.run.aoi_index)
    # the source follows the distributions: vz ~ N(184, 16), positions inside the 2 cm source disc
    assert abs(V[2].mean() - vdist.vz) < 0.2 and abs(V[2].std() - vdist.sigmaz) < 0.2
    assert np.all(X[0] ** 2 + X[1] ** 2 <= (xdist.d / 2) ** 2 * (1 + 1e-15)) and np.all(X[2] == xdist.z)


def test_device_drawn_run_with_lens_states_equals_the_host_arrays_path(ctx):
    """simulate_device with per-molecule lens states (config 4's four tables): the same tallies per
    state and the same kept molecules as simulate_arrays on the same draws and states."""
    from trajectories.trajectory_simulator import device_stream_keys, simulate_arrays, simulate_device

    bl = configs.config4()
    xdist, vdist = configs.source()
    n = 120_000
    states = np.random.RandomState(8).randint(0, 4, n).astype(np.uint8)
    keep = ["Detected", "Inside lens"]
    dev = simulate_device(bl, xdist, vdist, n, 5, keep, ctx=ctx, chunk=50_000, states=states)
    kx, kv = device_stream_keys(5)
    X = np.array([a.to_host()[:n] for a in xdist.draw_device(n, kx, ctx=ctx)])
    V = np.array([a.to_host()[:n] for a in vdist.draw_device(n, kv, ctx=ctx)])
    host = simulate_arrays(bl, X, V, keep, ctx=ctx, states=states)
    np.testing.assert_array_equal(dev.tally, host.tally)
    assert np.count_nonzero(dev.tally.sum(axis=1)) == 4
    np.testing.assert_array_equal(dev.aoi_index, host.aoi_index)
    np.testing.assert_array_equal(dev.molecules._init, host.molecules._init)
    np.testing.assert_array_equal(dev.molecules._states, host.molecules._states)
    for k in dev.aoi_final:
        np.testing.assert_array_equal(dev.aoi_final[k], host.aoi_final[k])
