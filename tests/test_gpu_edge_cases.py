"""GPU: the reference's failure behaviour and unusual inputs, through the C ABI.

* A reference-mode lens (RK4 variant, a(r) table to 1.01 d/2) whose stage points leave the table:
  the reference's interp1d raises ValueError there (electrostatic_lens.py:195-197,209,217); here every
  walker flags the same molecules as the oracle (CTRAJ_ST_OOT), run_simulation raises ValueError,
  and raise_out_of_table=False returns the oracle's counts.
* ctraj_set_beamline / ctraj_set_rtable / ctraj_set_wspline / ctraj_propagate / ctraj_sample_rows
  reject invalid tables and arguments with CTRAJ_EINVAL before any launch, and the context stays usable.
* Molecules with NaN or infinite positions, and with negative, zero or NaN v_z, are ordinary inputs:
  every walker gives the oracle's bits.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from trajectories import _native as N
from trajectories import configs
from trajectories.beamline import Beamline
from trajectories.beamline_elements import CircularAperture, ElectrostaticLens
from trajectories.compile import compile_elements
from trajectories.distributions import CeNTREXVelocityDistribution, GaussianPositionDistribution
from trajectories.engine import Engine
from trajectories.stark import TwoLevelStark

from beamlines import reference_example_beamline

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "t")
WALKERS = ((0, "two phases"), (N.RUN_LANES, "lane walker"), (N.RUN_SIMPLE, "one molecule per thread"))


def oot_beamline():
    """A reference-mode lens close to the source (its a(r) table ends at 1.01 d/2) and a detector."""
    lens = ElectrostaticLens(z0=0.05, L=0.3, name="ES lens", state=TwoLevelStark(sign=+1.0), integrator="ref")
    return Beamline([lens, CircularAperture(z0=0.6, L=0.0, d=0.01, name="Detector")])


def oot_source(n, seed):
    """Molecules spread over the bore with transverse speeds of tens of m/s: within one 1 mm step
    (5.4 us) some move more than 1 % of the bore radius outward, so a stage point leaves the table."""
    rs = np.random.RandomState(seed)
    V = np.array([rs.normal(0, 60.0, n), rs.normal(0, 60.0, n), rs.normal(184.0, 16.0, n)])
    X = np.array([rs.normal(0, 4e-3, n), rs.normal(0, 4e-3, n), np.zeros(n)])
    return X, V


def test_out_of_table_flags_equal_the_oracle_on_every_walker():
    bl = oot_beamline()
    X, V = oot_source(40_000, 1)
    eng = Engine(bl)
    orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2])
    n_oot = int(np.count_nonzero(orc["status"] & N.ST_OOT))
    assert n_oot > 50, n_oot  # the case is exercised
    for flags, name in WALKERS:
        dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], flags=flags)
        np.testing.assert_array_equal(dev["status"], orc["status"], err_msg=name)
        np.testing.assert_array_equal(dev["label"], orc["label"], err_msg=name)
        assert int(dev["work"][N.WORK_OOT]) == n_oot, name
        for k in FIELDS:
            np.testing.assert_array_equal(dev["out"][k], orc["out"][k], err_msg=f"{name} {k}")


def test_run_simulation_raises_like_interp1d_and_counts_when_asked_not_to():
    from trajectories.trajectory_simulator import TrajectorySimulator, reference_chunking

    bl = oot_beamline()
    vdist = CeNTREXVelocityDistribution(sigmax=60.0, sigmay=60.0)
    xdist = GaussianPositionDistribution(sigmax=4e-3, sigmay=4e-3, z=0.0)
    sim = TrajectorySimulator()
    np.random.seed(17)
    with pytest.raises(ValueError, match="outside its table"):
        sim.run_simulation(bl, "oot", vdist=vdist, xdist=xdist, N_traj=20_000, n_jobs=1)
    np.random.seed(17)
    sim.run_simulation(bl, "oot", vdist=vdist, xdist=xdist, N_traj=20_000, n_jobs=1, raise_out_of_table=False)
    assert sim.result.run.n_out_of_table > 0
    # the same draws through the oracle
    np.random.seed(17)
    loops, n_chunk, total = reference_chunking(20_000, 1)
    vs, xs = [], []
    for _ in range(loops):
        vs.append(vdist.draw(n_chunk))
        xs.append(xdist.draw(n_chunk))
    Vv, Xx = np.concatenate(vs, axis=1), np.concatenate(xs, axis=1)
    cb = compile_elements(bl.elements, bl.g)
    orc = O.propagate(cb, Xx[0], Xx[1], Xx[2], Vv[0], Vv[1], Vv[2])
    want = {}
    for lab in orc["label"]:
        want[cb.labels[lab]] = want.get(cb.labels[lab], 0) + 1
    assert sim.counter.counter_dict == want
    assert list(sim.counter.counter_dict) == list(want)
    assert sim.result.run.n_out_of_table == int(np.count_nonzero(orc["status"] & N.ST_OOT))


def _set(ctx, elems, n, phys):
    return N.lib().ctraj_set_beamline(ctx.handle, elems, n, C.byref(phys))


def _copy_elems(cb):
    arr = (N.Elem * cb.n_elems)()
    C.memmove(arr, cb.elems, C.sizeof(N.Elem) * cb.n_elems)
    return arr


def _copy_phys(cb):
    p = N.Phys()
    C.memmove(C.byref(p), C.byref(cb.phys), C.sizeof(N.Phys))
    return p


def test_invalid_tables_and_arguments_return_einval_before_any_launch(ctx):
    """check_elems / check_phys (ctraj.hip) and the table and run checks: each invalid input returns
    CTRAJ_EINVAL with a message, the beamline set before stays in force, and a run afterwards equals
    the oracle."""
    L = N.lib()
    eng = Engine(configs.config1(), ctx)
    cb = eng.cb
    lens_i = [e.kind for e in cb.elems[: cb.n_elems]].index(N.LENS)

    def expect_einval(rc, what):
        assert rc == N.EINVAL, f"{what}: rc={rc}"
        assert L.ctraj_last_error(ctx.handle), what

    e = _copy_elems(cb)
    e[0].z0, e[1].z0 = e[1].z0, e[0].z0  # unsorted z0
    expect_einval(_set(ctx, e, cb.n_elems, _copy_phys(cb)), "unsorted z0")
    e = _copy_elems(cb)
    e[0].label_hit = cb.phys.n_labels
    expect_einval(_set(ctx, e, cb.n_elems, _copy_phys(cb)), "label_hit >= n_labels")
    e = _copy_elems(cb)
    e[0].label_hit = -1
    expect_einval(_set(ctx, e, cb.n_elems, _copy_phys(cb)), "label_hit < 0")
    e = _copy_elems(cb)
    e[lens_i].label_inside = cb.phys.n_labels + 3
    expect_einval(_set(ctx, e, cb.n_elems, _copy_phys(cb)), "label_inside out of range")
    for dt in (0.0, -1e-5, float("nan")):
        e = _copy_elems(cb)
        e[lens_i].p[2] = dt
        expect_einval(_set(ctx, e, cb.n_elems, _copy_phys(cb)), f"RK4 lens dt={dt}")
    e = _copy_elems(cb)
    e[0].kind = 9
    expect_einval(_set(ctx, e, cb.n_elems, _copy_phys(cb)), "unknown kind")
    e = _copy_elems(cb)
    e[lens_i].force = 7
    expect_einval(_set(ctx, e, cb.n_elems, _copy_phys(cb)), "bad force model")
    for field, bad in (("label_detected", -1), ("label_detected", 64), ("n_labels", 0), ("n_states", 0),
                       ("n_states", 5), ("max_lens_steps", -1), ("max_lens_steps", (1 << 30) + 1)):
        p = _copy_phys(cb)
        setattr(p, field, bad)
        expect_einval(_set(ctx, _copy_elems(cb), cb.n_elems, p), f"phys.{field}={bad}")
    expect_einval(_set(ctx, _copy_elems(cb), N.MAX_ELEMS + 1, _copy_phys(cb)), "n_elems > 64")
    # tables
    r = np.array([0.0, 1e-3, 1e-3, 2e-3])
    a = np.zeros(4)
    expect_einval(L.ctraj_set_rtable(ctx.handle, 0, r.ctypes.data, a.ctypes.data, 4), "r knots not increasing")
    r2 = np.array([0.0, 1e-3, 2e-3])
    a2 = np.array([0.0, np.nan, 1.0])
    expect_einval(L.ctraj_set_rtable(ctx.handle, 0, r2.ctypes.data, a2.ctypes.data, 3), "NaN table entry")
    expect_einval(L.ctraj_set_rtable(ctx.handle, 9, r2.ctypes.data, r2.ctypes.data, 3), "table id 9")
    coef = np.zeros((1, 8, 4))
    coef[0, 3, 1] = np.inf
    expect_einval(L.ctraj_set_wspline(ctx.handle, 1, 8, coef.ctypes.data), "non-finite spline")
    # runs
    R = N.Run()
    R.n = -1
    expect_einval(L.ctraj_propagate(ctx.handle, C.byref(R)), "n < 0")
    R.n = 10
    expect_einval(L.ctraj_propagate(ctx.handle, C.byref(R)), "missing inputs")
    S = N.Sample()
    S.n = 2
    off = ctx.to_device(np.array([0, 5, 3], np.int64))
    rows = ctx.alloc(50)
    S.row_off, S.n_rows, S.rows, S.z = off.ptr, 5, rows.ptr, 0.1
    expect_einval(L.ctraj_sample_rows(ctx.handle, C.byref(S)), "decreasing row offsets")
    off2 = ctx.to_device(np.array([0, 3, 6], np.int64))
    S.row_off = off2.ptr
    expect_einval(L.ctraj_sample_rows(ctx.handle, C.byref(S)), "row offsets past n_rows")
    xs, ys = ctx.alloc(4), ctx.alloc(4)
    for radius in (-1.0, float("nan"), float("inf")):
        expect_einval(L.ctraj_draw_disc(ctx.handle, 1, 0, 4, radius, 0.0, xs.ptr, ys.ptr, None), f"disc radius {radius}")
    expect_einval(L.ctraj_draw_disc(ctx.handle, 1, 0, -1, 0.01, 0.0, xs.ptr, ys.ptr, None), "disc n < 0")
    expect_einval(L.ctraj_draw_disc(ctx.handle, 1, 0, 4, 0.01, 0.0, None, ys.ptr, None), "disc without x")
    # the valid beamline set before is still in force
    ctx.bound = None
    X, V = configs.draw_source(3000, seed=61)
    dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2])
    orc = O.propagate(cb, X[0], X[1], X[2], V[0], V[1], V[2])
    np.testing.assert_array_equal(dev["label"], orc["label"])
    # a run with a bad block size is refused too
    with pytest.raises(N.NativeError):
        eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], block=100)


def special_inputs(X, V):
    """Overwrite a spread of molecules with non-finite positions and non-positive or NaN v_z."""
    X, V = X.copy(), V.copy()
    n = X.shape[1]
    k = np.arange(0, n, 13)
    kinds = [
        (0, 0, np.nan), (0, 1, np.nan), (0, 0, np.inf), (0, 1, -np.inf), (0, 2, np.nan),
        (1, 2, -184.0), (1, 2, -1.0), (1, 2, 0.0), (1, 2, -0.0), (1, 2, np.nan), (1, 0, np.nan), (1, 1, np.inf),
    ]
    for j, i in enumerate(k):
        which, f, v = kinds[j % len(kinds)]
        (X if which == 0 else V)[f, i] = v
    return X, V


@pytest.mark.parametrize("case", ["config1", "config0", "reference example", "config4"])
def test_nan_and_nonpositive_vz_inputs_equal_the_oracle(case):
    if case == "config1":
        bl, ns = configs.config1(), None
    elif case == "config0":
        bl, ns = configs.config0(), None
    elif case == "config4":
        bl, ns = configs.config4(), 4
    else:
        bl, ns = reference_example_beamline(), None
    X, V = configs.draw_source(20_000, seed=71)
    if case == "reference example":
        X[:2] *= 0.05
        V[:2] *= 0.05
    X, V = special_inputs(X, V)
    st = None if ns is None else np.random.RandomState(72).randint(0, ns, X.shape[1]).astype(np.uint8)
    eng = Engine(bl, n_states=ns, max_lens_steps=4096)
    orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2], state=st)
    walkers = WALKERS if any(e.kind == N.LENS for e in eng.cb.elems[: eng.cb.n_elems]) else WALKERS[2:]
    for flags, name in walkers:
        dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], state=st, flags=flags, want_accel=True)
        np.testing.assert_array_equal(dev["label"], orc["label"], err_msg=name)
        np.testing.assert_array_equal(dev["status"], orc["status"], err_msg=name)
        for k in FIELDS:
            np.testing.assert_array_equal(dev["out"][k], orc["out"][k], err_msg=f"{name} {k}")
        for q in range(3):
            np.testing.assert_array_equal(dev["out_a"][q], orc["out_a"][q], err_msg=f"{name} a{q}")
        np.testing.assert_array_equal(dev["tally"].sum(axis=0), np.bincount(orc["label"], minlength=len(eng.labels)))
        # the oracle counts faces and steps; the device also counts flagged molecules
        np.testing.assert_array_equal(dev["work"][:2], orc["work"][:2], err_msg=name)
        assert int(dev["work"][N.WORK_OOT]) == int(np.count_nonzero(orc["status"] & N.ST_OOT)), name
        assert int(dev["work"][N.WORK_CAPPED]) == int(np.count_nonzero(orc["status"] & N.ST_CAPPED)), name


def test_step_cap_is_flagged_and_warned():
    """Molecules too slow for max_lens_steps full RK4 steps get CTRAJ_ST_CAPPED, as in the oracle, and
    run_simulation warns about them."""
    from trajectories.trajectory_simulator import simulate_arrays

    bl = Beamline([configs.lens(), CircularAperture("Detector", 2.5, 0.0, d=8e-3)], configs.G)
    X, V = configs.draw_source(4000, seed=81)
    X[2, ::5] = 0.79  # start just before the lens, so that gravity does not stop them first
    V[2, ::5] = 2.0  # 0.6 m at 2 m/s: 30,000 steps of 10 us
    eng = Engine(bl, max_lens_steps=2000)
    orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2])
    assert np.any(orc["status"] & N.ST_CAPPED)
    for flags, name in WALKERS:
        dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], flags=flags)
        np.testing.assert_array_equal(dev["status"], orc["status"], err_msg=name)
        np.testing.assert_array_equal(dev["label"], orc["label"], err_msg=name)
        for k in FIELDS:
            np.testing.assert_array_equal(dev["out"][k], orc["out"][k], err_msg=f"{name} {k}")
    with pytest.warns(RuntimeWarning, match="max_lens_steps"):
        res = simulate_arrays(bl, X, V, engine=eng)
    assert res.n_capped == int(np.count_nonzero(orc["status"] & N.ST_CAPPED))


def test_phase_times_bracket_each_kernel_of_a_launch(ctx):
    """CTRAJ_RUN_PHASE_TIMES: every timed launch reports its kernels' times once, oldest first; a
    single-kernel run has all of its time in the lens entry."""
    eng = Engine(configs.config1(), ctx)
    X, V = configs.draw_source(100_000, seed=95)
    inp = {k: ctx.to_device(a) for k, a in zip(FIELDS[:6], (X[0], X[1], X[2], V[0], V[1], V[2]))}
    tally = ctx.alloc(eng.cb.n_labels, np.int64).fill(0)
    ctx.phase_times()
    for flags in (0, 0, N.RUN_LANES):
        eng.launch(inp, tally=tally, n=X.shape[1], flags=flags | N.RUN_PHASE_TIMES)
    eng.launch(inp, tally=tally, n=X.shape[1])  # not timed
    w, ln, po = ctx.phase_times()
    assert len(w) == len(ln) == len(po) == 3
    assert np.all(w[:2] > 0) and np.all(ln > 0) and np.all(po[:2] >= 0)
    assert w[2] == 0 and po[2] == 0
    assert all(len(a) == 0 for a in ctx.phase_times())
    np.testing.assert_array_equal(tally.to_host().sum(), 4 * X.shape[1])
