"""Pins of the CPU oracle (oracle/ctraj_oracle.c) -- CPU only.

Every check compares the oracle with something it does not share code with: closed forms,
exact rational arithmetic, physical invariants, convergence order, the reference's own smoke case
(apertures.py:289-306) and the per-molecule numpy restatement (oracle/pyref.py).  Each is chosen
so that a plausible slip -- a dropped term, a wrong sign, a half step in the wrong stage, a
transposed operand -- fails it.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import oracle as O
from oracle import pyref
from trajectories import configs
from trajectories.beamline import Beamline
from trajectories.beamline_elements import (CircularAperture, ElectrostaticLens, FieldPlates, LinearTable,
                                            RectangularAperture)
from trajectories.compile import compile_elements
from trajectories.stark import TwoLevelStark, WSpline

from beamlines import reference_example_beamline

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "t")
G_REF = 9.80665


def run(bl, X, V, **kw):
    cb = compile_elements(bl.elements, bl.g)
    return cb, O.propagate(cb, X[0], X[1], X[2], V[0], V[1], V[2], **kw)


def test_smoke_case_of_the_reference():
    """apertures.py:289-306: a d = 0 aperture at z = 0.99 m stops the default molecule x0 = 0,
    v0 = (0, 0, 200) because gravity has moved it below the axis."""
    bl = Beamline([CircularAperture(name="40K Shield", z0=0.99, L=0.01, d=0),
                   RectangularAperture(name="Detection aperture", z0=100, L=0.01, w=0.00001, h=1e-6)])
    X = np.zeros((3, 1))
    V = np.array([[0.0], [0.0], [200.0]])
    cb, r = run(bl, X, V, rows_cap=8)
    assert cb.labels[r["label"][0]] == "40K Shield"
    assert r["n_rows"][0] == 2
    dt = 0.99 / 200
    np.testing.assert_allclose(r["rows"][0, 1, 1], -G_REF * dt * dt / 2, rtol=1e-15)
    np.testing.assert_allclose(r["rows"][0, 1, 9], dt, rtol=1e-15)


def test_ballistic_walk_matches_closed_form():
    bl = configs.config0()
    X, V = configs.draw_source(20_000, seed=100)
    cb, r = run(bl, X, V)
    g = bl.g
    z_stop = np.array([e.z0 for e in bl.elements] + [bl.elements[-1].z1])
    names = cb.labels
    for i in range(0, 20_000, 7):
        lab = names[r["label"][i]]
        zf = 2.0 if lab in ("Detected", "Detector") else [e.z0 for e in bl.elements if e.name == lab][0]
        t = (zf - X[2, i]) / V[2, i]
        xe = X[0, i] + V[0, i] * t
        ye = X[1, i] + V[1, i] * t - g * t * t / 2
        assert r["out"]["z"][i] == pytest.approx(zf, abs=1e-12)
        assert r["out"]["x"][i] == pytest.approx(xe, rel=1e-11, abs=1e-15)
        assert r["out"]["y"][i] == pytest.approx(ye, rel=1e-11, abs=1e-15)
        assert r["out"]["vy"][i] == pytest.approx(V[1, i] - g * t, rel=1e-12, abs=1e-12)
        assert r["out"]["t"][i] == pytest.approx(t, rel=1e-12)
    # a molecule stopped at an aperture really is outside it there, a survivor inside all of them
    rho = np.hypot(r["out"]["x"], r["out"]["y"])
    for e in bl.elements:
        hit = r["label"] == names.index(e.name)
        assert np.all(rho[hit] > e.d / 2)
    det = r["label"] == names.index("Detected")
    assert np.all(rho[det] <= bl.elements[-1].d / 2)
    assert z_stop.size == 5


def test_edges_are_strict_as_in_the_reference():
    """CircularAperture stops iff rho > d/2 (on the edge passes); RectangularAperture passes iff
    x1 < x < x2 and y1 < y < y2 (on the edge stops).  g = 0 keeps positions exact."""
    X = np.array([[0.002, 0.002, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    V = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [100.0, 100.0, 100.0]])
    circ = Beamline([CircularAperture("c", 0.5, 0.0, d=0.004)], 0.0)
    cb, r = run(circ, X, V)
    assert [cb.labels[k] for k in r["label"]] == ["Detected", "Detected", "Detected"]
    circ2 = Beamline([CircularAperture("c", 0.5, 0.0, d=0.004 - 1e-12)], 0.0)
    cb, r = run(circ2, X, V)
    assert [cb.labels[k] for k in r["label"]] == ["c", "c", "Detected"]
    rect = Beamline([RectangularAperture("r", 0.5, 0.0, w=0.004, h=0.01)], 0.0)
    cb, r = run(rect, X, V)
    assert [cb.labels[k] for k in r["label"]] == ["r", "r", "Detected"]


def test_field_plates_wall_hit_time():
    """FieldPlates (apertures.py:227-270): a molecule leaving the gap stops on the plate it reaches."""
    bl = Beamline([FieldPlates("plates", 1.0, 1.0, w=0.02)], 0.0)
    X = np.array([[0.0, 0.0], [0.0, 0.0], [0.0, 0.0]])
    V = np.array([[5.0, -1.0], [0.0, 0.0], [100.0, 100.0]])
    cb, r = run(bl, X, V)
    # first: enters at x = 0.05 > 0.01 -> stopped at the entrance; second: x(z0) = -0.01 exactly
    assert cb.labels[r["label"][0]] == "plates" and r["out"]["z"][0] == pytest.approx(1.0)
    assert cb.labels[r["label"][1]] == "plates"
    V2 = np.array([[0.15], [0.0], [100.0]])  # x(1.0) = 0.0015, x(2.0) = 0.003; no wall
    cb, r = run(bl, np.zeros((3, 1)), V2)
    assert cb.labels[r["label"][0]] == "Detected" and r["out"]["z"][0] == pytest.approx(2.0)
    V3 = np.array([[0.6], [0.0], [100.0]])  # x(1.0) = 0.006 inside, x(2.0) = 0.012 outside: wall at x = 0.01
    cb, r = run(bl, np.zeros((3, 1)), V3)
    assert cb.labels[r["label"][0]] == "plates"
    assert r["out"]["x"][0] == pytest.approx(0.01, rel=1e-14)
    assert r["out"]["t"][0] == pytest.approx(0.01 / 0.6, rel=1e-14)


def _ref_step_matrix(dt: Fr, k: Fr):
    """Exact 2x2 map (x, v) -> (x', v') of one step of the reference's lens integrator
    (electrostatic_lens.py:91-107) for the linear force a = -k x, in rational arithmetic."""
    def step(x, v):
        k1 = v
        l1 = -k * x
        k2 = k1 + dt * l1 / 2
        l2 = -k * (x + dt * k1)
        k3 = k1 + dt * l2 / 2
        l3 = -k * (x + dt * k2 / 2)
        k4 = k1 + dt * l3
        l4 = -k * (x + dt * k3)
        return x + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6, k1 + dt * (l1 + 2 * l2 + 2 * l3 + l4) / 6
    c1 = step(Fr(1), Fr(0))
    c2 = step(Fr(0), Fr(1))
    return [[c1[0], c2[0]], [c1[1], c2[1]]]


def test_reference_lens_integrator_exact_for_linear_force():
    """With a linear a_r = -k r table and g = 0 the reference lens is N = rint(L/dz) applications
    of an exactly computable 2x2 map.  The oracle must reproduce it to rounding; the stage
    structure (second stage at the full Euler point) changes the map at O(k dt^2)."""
    k = 2.0e6  # 1/s^2
    R = 0.02
    r_knots = np.linspace(0.0, 1.01 * R, 201)
    lens = ElectrostaticLens("lens", 0.1, 0.05, d=2 * R, dz=1e-3, a_interp=LinearTable(r_knots, -k * r_knots))
    bl = Beamline([lens], 0.0)
    vz = 125.0  # dt = 8 us exactly representable
    X = np.array([[1.0e-3], [-0.5e-3], [0.1]])
    V = np.array([[2.0], [1.0], [vz]])
    cb, r = run(bl, X, V)
    N = int(np.rint(0.05 / 1e-3))
    dt = Fr(1e-3) / Fr(vz)
    M = _ref_step_matrix(dt, Fr(k))
    for comp, (x0, v0) in enumerate(((X[0, 0], V[0, 0]), (X[1, 0], V[1, 0]))):
        x, v = Fr(x0), Fr(v0)
        for _ in range(N):
            x, v = M[0][0] * x + M[0][1] * v, M[1][0] * x + M[1][1] * v
        got_x = r["out"]["x" if comp == 0 else "y"][0]
        got_v = r["out"]["vx" if comp == 0 else "vy"][0]
        assert got_x == pytest.approx(float(x), rel=1e-12)
        assert got_v == pytest.approx(float(v), rel=1e-12)
    # the classical map (stage 2 at x + dt k1 / 2) differs from the reference's by far more than that
    kk = Fr(k)
    x, v = Fr(X[0, 0]), Fr(V[0, 0])
    for _ in range(N):
        k1, l1 = v, -kk * x
        k2, l2 = v + dt * l1 / 2, -kk * (x + dt * k1 / 2)
        k3, l3 = v + dt * l2 / 2, -kk * (x + dt * k2 / 2)
        k4, l4 = v + dt * l3, -kk * (x + dt * k3)
        x, v = x + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6, v + dt * (l1 + 2 * l2 + 2 * l3 + l4) / 6
    assert abs(float(x) - r["out"]["x"][0]) > 1e-6 * abs(float(x))


def _linear_w_spline(c):
    """Spline of W(E) = c E (exactly linear, so the cubic spline is exact and W' = c everywhere)."""
    E = np.linspace(0.0, 100e3, 1001) * 100.0
    return WSpline(E, c * E)


def _lens_only(spl, dt, V0=30e3, g=0.0, L=0.6):
    lens = ElectrostaticLens("lens", 0.0, L, d=0.035, V=V0, stark=spl, field="hexapole", integrator="rk4", dt=dt)
    return Beamline([lens], g)


@pytest.mark.parametrize("sign", [+1.0, -1.0])
def test_rk4_lens_harmonic_closed_form_and_order(sign):
    """W = c E with |E| = cE r^2 gives a harmonic transverse force -+ w^2 r (w^2 = 2 c cE / m).
    The lens output must match the closed form and converge at 4th order in dt."""
    c = sign * 2.0e-30  # J per V/m
    spl = _linear_w_spline(c)
    m = ElectrostaticLens("l", 0, 0.6).mass
    cE = 3 * 30e3 / 0.0175 ** 3
    w2 = 2 * c * cE / m  # a = -(2 c cE / m) r  ->  r'' = -w2 r
    X = np.array([[1.0e-3], [-2.0e-3], [0.0]])
    V = np.array([[0.3], [-0.2], [170.0]])
    T = 0.6 / 170.0
    g = 9.81

    def exact(x0, v0, gy):
        if w2 > 0:
            w = np.sqrt(w2)
            xs = gy / w2
            return (xs + (x0 - xs) * np.cos(w * T) + v0 / w * np.sin(w * T),
                    -(x0 - xs) * w * np.sin(w * T) + v0 * np.cos(w * T))
        w = np.sqrt(-w2)
        xs = -gy / w ** 2
        return (xs + (x0 - xs) * np.cosh(w * T) + v0 / w * np.sinh(w * T),
                (x0 - xs) * w * np.sinh(w * T) + v0 * np.cosh(w * T))

    errs = []
    for dt in (40e-6, 20e-6, 10e-6):
        cb, r = run(_lens_only(spl, dt, g=g), X, V)
        assert cb.labels[r["label"][0]] == "Detected"
        ex, evx = exact(X[0, 0], V[0, 0], 0.0)
        ey, evy = exact(X[1, 0], V[1, 0], -g)
        assert r["out"]["z"][0] == 0.6
        assert r["out"]["t"][0] == pytest.approx(T, rel=1e-13)
        errs.append(max(abs(r["out"]["x"][0] - ex) / abs(ex), abs(r["out"]["y"][0] - ey) / abs(ey),
                        abs(r["out"]["vx"][0] - evx) / abs(evx)))
    assert errs[-1] < 1e-8
    ratios = [errs[i] / errs[i + 1] for i in range(2)]
    for q in ratios:
        assert 11.0 < q < 22.0, (errs, ratios)  # 2^4 = 16 for a 4th-order method


def test_zero_stark_shift_lens_is_a_ballistic_drift():
    spl = _linear_w_spline(0.0)
    X, V = configs.draw_source(200, seed=7)
    bl = _lens_only(spl, 10e-6, g=9.81, L=0.6)
    cb, r = run(bl, X, V)
    t = (0.6 - X[2]) / V[2]
    np.testing.assert_allclose(r["out"]["x"], X[0] + V[0] * t, rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(r["out"]["y"], X[1] + V[1] * t - 9.81 * t * t / 2, rtol=1e-11, atol=1e-16)
    np.testing.assert_allclose(r["out"]["vy"], V[1] - 9.81 * t, rtol=1e-12)


def test_energy_and_angular_momentum_in_the_lens():
    """Transverse energy 1/2 m v^2 + W(|E|) + m g y is conserved (to O(dt^4)); with g = 0 the force
    is central and L_z = x vy - y vx is conserved."""
    stark = TwoLevelStark(sign=+1.0)
    spl = WSpline.from_stark(stark)
    m = ElectrostaticLens("l", 0, 0.6).mass
    cE = 3 * 30e3 / 0.0175 ** 3
    X, V = configs.draw_source(300, seed=9)
    for g in (9.81, 0.0):
        bl = _lens_only(spl, 10e-6, g=g)
        cb, r = run(bl, X, V)
        ok = r["label"] == cb.labels.index("Detected")
        assert ok.sum() > 100
        x, y, vx, vy = (r["out"][k][ok] for k in ("x", "y", "vx", "vy"))
        x0, y0, vx0, vy0 = X[0, ok], X[1, ok], V[0, ok], V[1, ok]
        def energy(x, y, vx, vy):
            return 0.5 * m * (vx ** 2 + vy ** 2) + spl.W_at(cE * (x * x + y * y)) + m * g * y
        e0, e1 = energy(x0, y0, vx0, vy0), energy(x, y, vx, vy)
        scale = 0.5 * m * (vx0 ** 2 + vy0 ** 2) + np.abs(spl.W_at(cE * (x0 ** 2 + y0 ** 2))) + 1e-40
        assert np.max(np.abs(e1 - e0) / scale) < 1e-7
        if g == 0.0:
            l0, l1 = x0 * vy0 - y0 * vx0, x * vy - y * vx
            assert np.max(np.abs(l1 - l0) / (np.abs(l0) + 1e-12)) < 1e-8


def test_zgrid_resampling_of_a_ballistic_molecule():
    """find_radial_pos_dist's rule on a ballistic trajectory gives the closed-form position at each z."""
    bl = Beamline([CircularAperture("a", 1.0, 0.0, d=1.0), CircularAperture("b", 2.5, 0.0, d=1.0)], 9.81)
    X = np.array([[1e-3], [2e-3], [0.0]])
    V = np.array([[1.0], [-0.5], [180.0]])
    cb, r = run(bl, X, V, rows_cap=16)
    K = 2501
    s = O.zgrid_from_rows(r["rows"][0], r["n_rows"][0], 0.0, 1e-3, K)
    z = np.arange(K) * 1e-3
    t = z / 180.0
    np.testing.assert_allclose(s[0], 1e-3 + 1.0 * t, rtol=1e-12)
    np.testing.assert_allclose(s[1], 2e-3 - 0.5 * t - 9.81 * t * t / 2, rtol=1e-10, atol=1e-18)
    np.testing.assert_allclose(s[3], -0.5 - 9.81 * t, rtol=1e-12)
    np.testing.assert_allclose(s[5], t, rtol=1e-12, atol=1e-18)


def _row(z, x=0.0, y=0.0, vx=0.0, vy=0.0, vz=100.0, ax=0.0, ay=-9.81, t=0.0):
    return [x, y, z, vx, vy, vz, ax, 0.0 if ay is None else ay, 0.0, t]


def test_resampling_rule_special_cases():
    """post_processing.py:40-75 on hand-built rows whose answers are known without the rule's code:
    the exact-z row is returned as stored even when it is off the parabola; of two rows at that z the
    first wins; the extrapolated row is the one before the FIRST row with z >= z_q (not the last row
    below z_q, which differs when z goes back); a z_q before the first row uses the last row (Python's
    index -1); a z_q beyond the last row gives nothing."""
    rows = np.array([
        _row(0.0, x=1e-3, vx=2.0, t=0.0),
        _row(0.4, x=5e-3, vx=2.0, t=0.004),
        _row(0.4, x=9e-3, vx=7.0, t=0.005),   # a second row at z = 0.4 (an L = 0 face): not chosen
        _row(0.2, x=-3e-3, vx=-1.0, t=0.006),  # z goes back
        _row(1.0, x=4e-3, y=-2e-3, vx=0.5, vy=-0.25, vz=200.0, t=0.01),
    ])
    n = rows.shape[0]

    def at(zq):
        return O.zgrid_from_rows(rows, n, zq, 1.0, 1)[:, 0]

    # exact hit: the first row at 0.4, as stored
    np.testing.assert_array_equal(at(0.4), [5e-3, 0.0, 2.0, 0.0, 100.0, 0.004])
    # z_q = 0.3: the first row with z >= 0.3 is row 1, so row 0 is advanced by dt = 0.3/100
    dt = 0.3 / 100.0
    np.testing.assert_allclose(at(0.3), [1e-3 + 2.0 * dt, -9.81 * dt * dt / 2, 2.0, -9.81 * dt, 100.0, dt],
                               rtol=1e-13, atol=1e-18)
    # z_q = 0.7: rows 0..3 are below, row 4 is the first >= 0.7, so row 3 (z = 0.2) is advanced
    dt = (0.7 - 0.2) / 100.0
    np.testing.assert_allclose(at(0.7), [-3e-3 - 1.0 * dt, -9.81 * dt * dt / 2, -1.0, -9.81 * dt, 100.0,
                                         0.006 + dt], rtol=1e-13, atol=1e-18)
    # z_q before the first row: index -1, the last row advanced backwards
    dt = (-0.5 - 1.0) / 200.0
    np.testing.assert_allclose(at(-0.5), [4e-3 + 0.5 * dt, -2e-3 - 0.25 * dt - 9.81 * dt * dt / 2, 0.5,
                                          -0.25 - 9.81 * dt, 200.0, 0.01 + dt], rtol=1e-13, atol=1e-18)
    # beyond the last row: skipped
    assert np.all(np.isnan(at(1.0 + 1e-9)))
    # exactly at the last row's z: returned as stored
    np.testing.assert_array_equal(at(1.0), [4e-3, -2e-3, 0.5, -0.25, 200.0, 0.01])


def test_philox_known_answers():
    """Published Philox4x32-10 known-answer vectors (Random123 kat_vectors)."""
    kats = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kats:
        got = O.philox(np.array(ctr, dtype=np.uint32), key)
        assert tuple(int(v) for v in got) == want


def test_gaussian_source_moments():
    outs = O.draw_gaussian(200_000, [0.0, 184.0], [1.0, 16.0], seed=42)
    z = outs[0]
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert abs(np.mean(z ** 3)) < 0.02 and abs(np.mean(z ** 4) - 3) < 0.05
    assert abs(outs[1].mean() - 184.0) < 0.1


def test_disc_source_from_its_philox_words_and_uniform():
    """The oracle's uniform-disc source (ctraj_draw_disc's restatement): each point is theta = 2 pi u1,
    r = R sqrt(u2) from the Philox words of counter (i, 0, 'DISC', 'CTRA'), and the points are uniform
    over the disc (r^2/R^2 and theta/2pi uniform, independent)."""
    R, z0, seed, off = 0.01, 0.00635, 0x1234_5678_9ABC_DEF0, 5
    x, y, z = O.draw_disc(200_000, R, z0, seed, offset=off)
    assert np.all(z == z0)
    for i in (0, 1, 77, 199_999):
        c = O.philox(np.array([i + off, 0, 0x43534944, 0x43545241], dtype=np.uint32),
                     (seed & 0xFFFFFFFF, seed >> 32)).astype(np.uint64)
        u1 = (float((c[0] << np.uint64(21)) | (c[1] >> np.uint64(11))) + 0.5) * 2.0 ** -53
        u2 = (float((c[2] << np.uint64(21)) | (c[3] >> np.uint64(11))) + 0.5) * 2.0 ** -53
        r, th = np.sqrt(u2) * R, 2 * np.pi * u1
        assert x[i] == pytest.approx(r * np.cos(th), rel=1e-14, abs=1e-18)
        assert y[i] == pytest.approx(r * np.sin(th), rel=1e-14, abs=1e-18)
    q = (x * x + y * y) / R ** 2
    a = (np.arctan2(y, x) / (2 * np.pi)) % 1.0
    assert np.all(q <= 1.0)
    for u in (q, a):
        assert abs(u.mean() - 0.5) < 0.003 and abs(u.var() - 1 / 12) < 0.002
    assert abs(np.corrcoef(q, a)[0, 1]) < 0.01
    # the offset shifts the counter: molecule i of offset k is molecule i + k of offset 0
    x0, y0, _ = O.draw_disc(10, R, z0, seed, offset=off + 3)
    np.testing.assert_array_equal(x0[:5], x[3:8])


def test_pyref_matches_oracle_reference_mode():
    """The numpy restatement (the reference's arithmetic) and the C oracle agree to rounding on a
    reference-mode beamline (apertures, reference lens with its a(r) table, field plates)."""
    bl = reference_example_beamline()
    rs = np.random.RandomState(3)
    n = 200
    V = np.array([rs.normal(0, 1.0, n), rs.normal(0, 1.0, n), rs.normal(184, 16, n)])
    X = np.array([rs.normal(0, 1e-3, n), rs.normal(0, 1e-3, n), np.full(n, 0.00635)])
    cb, r = run(bl, X, V)
    hits, fin = pyref.propagate_block(bl.elements, bl.g, X, V)
    lab = np.array([cb.labels.index(h) for h in hits])
    np.testing.assert_array_equal(lab, r["label"])
    o = np.stack([r["out"][k] for k in FIELDS], 1)
    np.testing.assert_allclose(fin, o, rtol=1e-12, atol=1e-15)
    assert len(set(hits)) >= 3


def test_pyref_matches_oracle_config_mode():
    for bl in (configs.config1(), configs.config1(stark=configs.stark_table(+1.0))):
        X, V = configs.draw_source(200, seed=12)
        cb, r = run(bl, X, V)
        hits, fin = pyref.propagate_block(bl.elements, bl.g, X, V)
        lab = np.array([cb.labels.index(h) for h in hits])
        np.testing.assert_array_equal(lab, r["label"])
        o = np.stack([r["out"][k] for k in FIELDS], 1)
        np.testing.assert_allclose(fin, o, rtol=1e-11, atol=1e-15)


def test_record_rows_follow_the_reference_row_rule():
    """Rows: initial point, every face, lens entrance, every lens step, the exit hop (molecule.py:
    115-144, electrostatic_lens.py:106-111)."""
    bl = reference_example_beamline()
    X = np.array([[0.0], [0.0], [0.00635]])
    V = np.array([[0.01], [0.01], [184.0]])
    cb, r = run(bl, X, V, rows_cap=800)
    n = r["n_rows"][0]
    lens = [e for e in bl.elements if isinstance(e, ElectrostaticLens)][0]
    N = int(np.rint(lens.L / lens.dz))
    # 3 circular apertures x 2 faces, lens entrance + N steps + exit hop, plates 2 rows, DR 2 faces
    assert cb.labels[r["label"][0]] == "Detected"
    assert n == 1 + 6 + (1 + N + 1) + 2 + 2
    assert np.all(np.diff(r["rows"][0, :n, 2]) >= -1e-12)


# ---------------------------------------------------------------------------------------------
# Smoothed detected flux and its reverse-mode gradient (orc_smooth)
# ---------------------------------------------------------------------------------------------
def _smooth_case(n, seed=3, sign=-1.0):
    bl = configs.config1(stark=configs.stark_table(sign))
    cb = compile_elements(bl.elements, bl.g)
    X, V = configs.draw_source(n, seed)
    eps = (V[2] - configs.VZ_MEAN) / configs.VZ_SIGMA
    theta = np.array([6e-3, 30e3, 0.8, 3e-3, 4e-3, configs.VZ_MEAN])
    return bl, cb, X, V, eps, theta


@pytest.mark.parametrize("softness", [50e-6, 1e-3])
def test_smooth_gradient_matches_richardson_central_differences(softness):
    """Reverse-mode gradient of the oracle's J against Richardson-extrapolated central differences
    of the oracle's own J, for every parameter (radii, lens V and z0, mu): a sign, a missing factor
    or a wrong stage weight in any adjoint term fails this.  Compared as theta_j dJ/dtheta_j; the
    wide sigmoid makes every aperture radius matter."""
    bl, cb, X, V, eps, theta = _smooth_case(400)
    kw = dict(sigma_vz=configs.VZ_SIGMA, softness=softness, max_steps=2000)
    r = O.smooth(cb, X[0], X[1], V[0], V[1], eps, theta, **kw)
    G = r["grad"][0]
    fd = []
    for j in range(theta.size):
        def d(hrel):
            h = hrel * abs(theta[j])
            P = np.repeat(theta[None], 2, axis=0)
            P[0, j] += h
            P[1, j] -= h
            J = O.smooth(cb, X[0], X[1], V[0], V[1], eps, P, grad=False, **kw)["J"]
            return (J[0] - J[1]) / (P[0, j] - P[1, j])
        fd.append((4 * d(1e-5) - d(2e-5)) / 3)
    gs, fs = G * theta, np.array(fd) * theta
    scale = np.max(np.abs(gs))
    if softness > 1e-4:
        assert np.all(np.abs(gs) > 1e-3 * scale)
    np.testing.assert_allclose(gs, fs, rtol=2e-6, atol=1e-9 * scale)


def test_smooth_sharp_limit_counts_detected_molecules():
    """With a vanishing sigmoid width J is the fraction of molecules inside every aperture: it must
    equal the detected fraction of the ordinary (hard-edged) walk with the lens bore opened up.
    (Focusing sign of the Stark shift, so that many molecules reach the detector.)"""
    bl, cb, X, V, eps, theta = _smooth_case(3000, seed=5, sign=+1.0)
    r = O.smooth(cb, X[0], X[1], V[0], V[1], eps, theta, sigma_vz=configs.VZ_SIGMA, softness=1e-15,
                 max_steps=2000, grad=False)
    cb.elems[1].p[0] = 1e9  # lens entrance radius
    cb.elems[1].p[1] = 1e18  # bore radius squared
    vz = configs.VZ_MEAN + configs.VZ_SIGMA * eps
    p = O.propagate(cb, X[0], X[1], X[2], V[0], V[1], vz)
    det = int(np.count_nonzero(p["label"] == cb.label_detected))
    assert det > 30
    assert r["J"][0] * 3000 == pytest.approx(det, abs=1e-9)


def test_smooth_parameter_sets_are_independent_and_mu_shifts_vz():
    """A batch of parameter sets equals the sets evaluated alone; mu enters only through vz."""
    bl, cb, X, V, eps, theta = _smooth_case(200, seed=7)
    kw = dict(sigma_vz=configs.VZ_SIGMA, softness=50e-6, max_steps=2000)
    P = np.stack([theta, theta * [1, 1.1, 1, 1, 1, 1], theta * [1, 1, 1, 1, 1, 0.9]])
    rb = O.smooth(cb, X[0], X[1], V[0], V[1], eps, P, **kw)
    for c in range(3):
        rs = O.smooth(cb, X[0], X[1], V[0], V[1], eps, P[c], **kw)
        assert rb["J"][c] == rs["J"][0]
        np.testing.assert_array_equal(rb["grad"][c], rs["grad"][0])
    # mu -> mu + d with eps -> eps - d/sigma describes the same molecules
    d = 2.0
    r2 = O.smooth(cb, X[0], X[1], V[0], V[1], eps - d / configs.VZ_SIGMA, theta + [0, 0, 0, 0, 0, d], **kw)
    assert r2["J"][0] == pytest.approx(rb["J"][0], rel=1e-12)
