"""GPU parity on random beamlines: every walker against the CPU oracle, bit for bit.

Each case draws a beamline from a fixed seed -- circular and rectangular apertures and field
plates at random positions and sizes, thin or with length, around one electrostatic lens in either
model (configuration mode: hexapole or linear field, W' spline of either Stark sign, random V, bore,
length and dt; reference mode: the a(r) table built from the two-level form, random dz) -- and a
source of random spread, then compares final phase space, stored acceleration, labels, status bits,
tallies and work counters of the default split path, the single-kernel lane walker and the
one-molecule-per-thread walker with the oracle; on every fourth seed it also compares the record
kernel's trajectory rows and z-grid samples.  Runs without per-molecule outputs (tallies, first
indices, work counters and kept molecules only) are compared on every seed and on a four-state
beamline.
"""
import numpy as np
import pytest

from oracle import oracle as O
from trajectories import _native as N
from trajectories.beamline import Beamline
from trajectories.beamline_elements import CircularAperture, ElectrostaticLens, FieldPlates, RectangularAperture
from trajectories.engine import Engine
from trajectories.stark import TwoLevelStark, WSpline

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "t")


def random_beamline(seed: int):
    rs = np.random.RandomState(seed)
    z = 0.05
    els = []

    def aperture(name):
        nonlocal z
        kind = rs.randint(3)
        L = 0.0 if rs.rand() < 0.5 else rs.uniform(1e-3, 0.05)
        if kind == 0:
            el = CircularAperture(name, z, L, x0=rs.uniform(-1e-3, 1e-3), y0=rs.uniform(-1e-3, 1e-3),
                                  d=rs.uniform(4e-3, 30e-3))
        elif kind == 1:
            el = RectangularAperture(name, z, L, x0=rs.uniform(-1e-3, 1e-3), y0=rs.uniform(-1e-3, 1e-3),
                                     w=rs.uniform(4e-3, 30e-3), h=rs.uniform(4e-3, 30e-3))
        else:
            el = FieldPlates(name, z, rs.uniform(0.05, 0.6), x0=rs.uniform(-1e-3, 1e-3), w=rs.uniform(6e-3, 30e-3))
        z = el.z0 + el.L + rs.uniform(0.01, 0.5)
        return el

    for k in range(rs.randint(0, 3)):
        els.append(aperture(f"before {k}"))
    ref = rs.rand() < 0.4
    L = rs.uniform(0.2, 0.8)
    if ref:
        lens = ElectrostaticLens("ES lens", z, L, d=rs.uniform(0.02, 0.05), V=rs.uniform(5e3, 40e3),
                                 dz=rs.uniform(0.5e-3, 2e-3), state=TwoLevelStark(sign=rs.choice([-1.0, 1.0])))
    else:
        stark = WSpline.from_stark(TwoLevelStark(sign=rs.choice([-1.0, 1.0]), dipole_scale=rs.uniform(0.5, 1.0)),
                                   100e3, 1001)
        lens = ElectrostaticLens("ES lens", z, L, d=rs.uniform(0.02, 0.05), V=rs.uniform(5e3, 40e3), stark=stark,
                                 field=rs.choice(["hexapole", "linear"]), integrator="rk4",
                                 dt=rs.uniform(5e-6, 20e-6))
    els.append(lens)
    z = lens.z1 + rs.uniform(0.05, 0.5)
    for k in range(rs.randint(1, 4)):
        els.append(aperture(f"after {k}"))
    order = rs.permutation(len(els))  # the Beamline sorts them by z0 (stable)
    bl = Beamline([els[i] for i in order], 9.80665)
    n = 12_000 if ref else 20_000
    sx, sv = rs.uniform(1e-3, 5e-3), rs.uniform(1.0, 20.0)
    X = np.vstack([rs.normal(0, sx, n), rs.normal(0, sx, n), np.zeros(n)])
    V = np.vstack([rs.normal(0, sv, n), rs.normal(0, sv, n), rs.normal(184.0, 20.0, n)])
    return bl, X, V, ref


@pytest.mark.parametrize("seed", list(range(40)))
def test_random_beamline_every_walker_equals_the_oracle(seed):
    bl, X, V, ref = random_beamline(1000 + seed)
    eng = Engine(bl)
    orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2])
    labels = np.bincount(orc["label"], minlength=len(eng.labels))
    assert labels[eng.cb.label_detected] + labels[eng.labels.index("Inside lens")] > 0, "the lens is reached"
    for flags in (0, N.RUN_LANES, N.RUN_SIMPLE, N.RUN_QUEUE_ORDER):
        dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], want_accel=True, flags=flags)
        tag = f"seed {seed} ({'reference' if ref else 'configuration'} mode) flags={flags}"
        np.testing.assert_array_equal(dev["label"], orc["label"], err_msg=tag)
        np.testing.assert_array_equal(dev["status"], orc["status"], err_msg=tag)
        for k in FIELDS:
            np.testing.assert_array_equal(dev["out"][k], orc["out"][k], err_msg=f"{tag} {k}")
        for q in range(3):
            np.testing.assert_array_equal(dev["out_a"][q], orc["out_a"][q], err_msg=f"{tag} a{q}")
        np.testing.assert_array_equal(dev["tally"][0], labels, err_msg=tag)
        assert int(dev["work"][N.WORK_FACES]) == int(orc["work"][0]), tag
        assert int(dev["work"][N.WORK_STEPS]) == int(orc["work"][1]), tag


@pytest.mark.parametrize("seed", list(range(0, 40, 4)))
def test_random_beamline_trajectory_rows_and_zgrid_equal_the_oracle(seed):
    """The record kernel on random beamlines: every trajectory row (the rows the reference's
    Trajectory would hold) and the 1-mm z-grid samples of the post-processing rule equal the
    oracle's, bit for bit."""
    bl, X, V, ref = random_beamline(1000 + seed)
    n = 800
    X, V = X[:, :n], V[:, :n]
    eng = Engine(bl)
    rows, nrows, lab = eng.record_rows(X[0], X[1], X[2], V[0], V[1], V[2], cap=96)  # grows to the longest
    orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2], rows_cap=rows.shape[1])
    np.testing.assert_array_equal(lab, orc["label"])
    np.testing.assert_array_equal(nrows, orc["n_rows"])
    np.testing.assert_array_equal(rows, orc["rows"])
    K = int(np.ceil(max(el.z1 for el in bl.elements) / 1e-3)) + 50
    samp, written, lab2 = eng.zgrid(X[0], X[1], X[2], V[0], V[1], V[2], z_start=0.0, z_step=1e-3, n_samples=K)
    np.testing.assert_array_equal(lab2, orc["label"])
    for i in range(0, n, 37):
        ref_s = O.zgrid_from_rows(orc["rows"][i], int(orc["n_rows"][i]), 0.0, 1e-3, K)
        np.testing.assert_array_equal(samp[:, :, i], ref_s, err_msg=f"seed {seed} molecule {i}")
        assert written[i] == np.count_nonzero(~np.isnan(ref_s[0]))


def check_tallies_only_run(eng, dev, orc, state, index_base, tag):
    """A run without per-molecule outputs, keeping only the detected molecules, against the oracle:
    tallies per state, first index per label (offset by index_base), work counters with the
    out-of-table and step-capped counts, and the kept molecules' indices and final phase space."""
    lab = orc["label"]
    nl = len(eng.labels)
    ns = dev["tally"].shape[0]
    for s in range(ns):
        m = state == s
        np.testing.assert_array_equal(dev["tally"][s], np.bincount(lab[m], minlength=nl), err_msg=f"{tag} state {s}")
    for i in range(nl):
        hit = np.nonzero(lab == i)[0]
        want = index_base + hit[0] if hit.size else np.iinfo(np.int64).max
        assert int(dev["first_index"][i]) == want, (tag, eng.labels[i])
    assert int(dev["work"][N.WORK_FACES]) == int(orc["work"][0]), tag
    assert int(dev["work"][N.WORK_STEPS]) == int(orc["work"][1]), tag
    assert int(dev["work"][N.WORK_OOT]) == int(np.count_nonzero(orc["status"] & 1)), tag
    assert int(dev["work"][N.WORK_CAPPED]) == int(np.count_nonzero(orc["status"] & 2)), tag
    det = np.nonzero(lab == eng.cb.label_detected)[0]
    o = np.argsort(dev["aoi_index"])
    np.testing.assert_array_equal(dev["aoi_index"][o], index_base + det, err_msg=tag)
    for k in FIELDS:
        np.testing.assert_array_equal(dev["aoi_out"][k][o], orc["out"][k][det], err_msg=f"{tag} kept {k}")


@pytest.mark.parametrize("seed", list(range(40)))
def test_random_beamline_tallies_only_run_equals_the_oracle(seed):
    """No per-molecule outputs: for a hexapole lens in RK4 steps the lens kernel tallies the
    molecules stopped in the lens itself and the walk after the lens reads only those that leave it;
    either way the tallies, first indices, work counters and kept molecules equal the oracle's."""
    bl, X, V, ref = random_beamline(1000 + seed)
    eng = Engine(bl)
    orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2])
    base = 7_000_000_000
    for flags in (0, N.RUN_QUEUE_ORDER):
        dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], outputs=False, per_molecule=False,
                            aoi_labels=("Detected",), index_base=base, flags=flags)
        check_tallies_only_run(eng, dev, orc, np.zeros(X.shape[1], np.uint8), base,
                               f"seed {seed} ({'reference' if ref else 'configuration'} mode) flags={flags}")


def test_four_state_tallies_only_run_equals_the_oracle():
    """The lens kernel's own tallies of stopped molecules are kept per lens state."""
    from trajectories import configs

    bl = configs.config4()
    n = 200_000
    X, V = configs.draw_source(n, seed=45)
    st = np.random.RandomState(46).randint(0, 4, n).astype(np.uint8)
    eng = Engine(bl, n_states=4)
    orc = O.propagate(eng.cb, X[0], X[1], X[2], V[0], V[1], V[2], state=st)
    dev = eng.propagate(X[0], X[1], X[2], V[0], V[1], V[2], state=st, outputs=False, per_molecule=False,
                        aoi_labels=("Detected",), index_base=123)
    assert dev["tally"].shape[0] == 4
    check_tallies_only_run(eng, dev, orc, st, 123, "config 4, four states")
