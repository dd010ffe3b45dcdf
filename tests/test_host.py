"""Host-side logic, CPU only: Stark tables and splines, source draws, beamline packing, Counter,
chunking, and the C-ABI library's exported symbols (loaded, never called into the GPU)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from trajectories import _native as N
from trajectories import configs
from trajectories.beamline import Beamline
from trajectories.beamline_elements import (CircularAperture, ElectrostaticLens, FieldPlates, LinearTable,
                                            RectangularAperture)
from trajectories.compile import compile_elements
from trajectories.distributions import (CeNTREXPositionDistribution, CeNTREXVelocityDistribution,
                                        GaussianPositionDistribution)
from trajectories.stark import DEBYE, H_PLANCK, TwoLevelStark, WSpline
from trajectories.trajectory_simulator import Counter, reference_chunking

ROOT = Path(__file__).resolve().parent.parent


def test_two_level_stark_closed_form():
    s = TwoLevelStark(sign=-1.0)
    E = np.array([0.0, 1e5, 5.14e6, 1e7])
    hnu2 = H_PLANCK * 6.67e9 / 2
    mu = 4.23 * DEBYE
    np.testing.assert_allclose(s.energy(E), -(np.sqrt(hnu2 ** 2 + (mu * E) ** 2) - hnu2), rtol=1e-15)
    assert s.energy(0.0) == 0.0
    # limits: quadratic at small fields, linear (slope -mu) at large fields
    assert s.energy(1.0) == pytest.approx(-(mu ** 2) / (2 * hnu2), rel=1e-6)
    assert s.denergy(1e9) == pytest.approx(-mu, rel=1e-6)
    # analytic derivative vs central difference
    h = 1.0
    assert s.denergy(3e6) == pytest.approx((s.energy(3e6 + h) - s.energy(3e6 - h)) / (2 * h), rel=1e-7)
    assert TwoLevelStark(sign=+1.0, dipole_scale=0.5).energy(2e6) == pytest.approx(
        np.sqrt(hnu2 ** 2 + (0.5 * mu * 2e6) ** 2) - hnu2, rel=1e-15)


@pytest.mark.parametrize("sign,scale", [(-1.0, 1.0), (1.0, 1.0), (-1.0, 0.5), (-1.0, 0.8)])
def test_wspline_matches_scipy_not_a_knot(sign, scale):
    spl = WSpline.from_stark(TwoLevelStark(sign=sign, dipole_scale=scale))
    ref = CubicSpline(spl.E, spl.W, bc_type="not-a-knot")
    # scipy's c[k] is the coefficient of (E - E_i)^(3-k)
    # the cubic coefficient is a difference of nearly equal second derivatives: compare on its scale
    np.testing.assert_allclose(spl.c3, ref.c[0], rtol=0, atol=1e-10 * np.abs(ref.c[0]).max())
    np.testing.assert_allclose(spl.c2, ref.c[1], rtol=1e-9, atol=1e-12 * np.abs(ref.c[1]).max())
    np.testing.assert_allclose(spl.c1, ref.c[2], rtol=1e-10, atol=1e-13 * np.abs(ref.c[2]).max())
    E = np.random.RandomState(0).uniform(0, 1e7, 5000)
    dref = ref(E, 1)
    np.testing.assert_allclose(spl.dW(E), dref, rtol=1e-9, atol=1e-12 * np.abs(dref).max())
    np.testing.assert_allclose(spl.W_at(E), ref(E), rtol=1e-9, atol=1e-14 * np.abs(spl.W).max())


def test_wspline_derivative_against_analytic_force_law():
    """W' from the spline vs the analytic dW/dE: the spline error bound (h^3 |W''''| scale) holds."""
    # The two-level W turns from quadratic to linear over E_c = h nu / (2 mu) ~ 1.6 kV/cm, within
    # ~15 knots of 0.1 kV/cm, where the derivative error of a cubic spline is largest; beyond it the
    # error falls to rounding level.
    s = TwoLevelStark(sign=-1.0)
    spl = WSpline.from_stark(s)
    E = np.linspace(0, 1e7, 20011)
    err = np.abs(spl.dW(E) - s.denergy(E)) / np.abs(s.denergy(1e7))
    assert err.max() < 5e-4
    assert err[E > 2e5].max() < 1e-5
    assert err[E > 1e6].max() < 1e-8
    # halving the knot spacing cuts the error by ~2^3 where the spline is resolved
    Ef = np.linspace(0, 1e7, 2001)
    fine = WSpline(Ef, s.energy(Ef))
    e2 = np.abs(fine.dW(E) - s.denergy(E)) / np.abs(s.denergy(1e7))
    sel = (E > 2e5) & (E < 1e6)
    assert 4.0 < err[sel].max() / e2[sel].max() < 20.0


def test_wspline_device_coefficients_and_clamp():
    spl = WSpline.from_stark(TwoLevelStark())
    assert spl.coef.shape == (1000, 4) and spl.dE == pytest.approx(1e4)
    E = np.array([0.5e4, 123456.0, 9.99e6])
    u = E / spl.dE
    k = u.astype(int)
    s = u - k
    A, B, C = spl.coef[k, 0], spl.coef[k, 1], spl.coef[k, 2]
    np.testing.assert_allclose((A * s + B) * s + C, spl.dW(E), rtol=1e-15)
    # beyond the table the force saturates at the last knot
    assert spl.dW(2e7) == pytest.approx(spl.dW(1e7), rel=1e-12)


def test_distribution_draws_replay_numpy_global_rng():
    """scipy.stats draws on numpy's global RandomState (distributions.py:40-41,69-76,155-162) are the
    RandomState normal stream: a seeded run replays the reference's inputs exactly."""
    xdist, vdist = configs.source()
    np.random.seed(5)
    v = vdist.draw(1000)
    x = xdist.draw(1000)
    rs = np.random.RandomState(5)
    np.testing.assert_array_equal(v[0], 0.0 + 1.5 * rs.standard_normal(1000))
    np.testing.assert_array_equal(v[1], 0.0 + 1.5 * rs.standard_normal(1000))
    np.testing.assert_array_equal(v[2], 184.0 + 16.0 * rs.standard_normal(1000))
    np.testing.assert_array_equal(x[0], 1.5e-3 * rs.standard_normal(1000))
    np.testing.assert_array_equal(x[1], 1.5e-3 * rs.standard_normal(1000))
    assert np.all(x[2] == 0.0) and x.shape == (3, 1000)
    X, V = configs.draw_source(1000 * 100, seed=5)
    np.random.seed(5)
    np.testing.assert_array_equal(V[:, :1000], vdist.draw(1000))
    np.testing.assert_array_equal(X[:, :1000], xdist.draw(1000))
    d = CeNTREXPositionDistribution().draw(2000)
    assert np.all(np.hypot(d[0], d[1]) <= 0.01) and np.all(d[2] == 0.25 * 0.0254)
    assert GaussianPositionDistribution().draw(10).shape == (3, 10)
    assert CeNTREXVelocityDistribution().draw(7).shape == (3, 7)


def test_reference_chunking():
    """run_simulation simulates int(N_traj/(100 n_jobs)) * 100 n_jobs molecules (trajectory_simulator.py:48-49)."""
    assert reference_chunking(100_000, 64) == (6400, 15, 96_000)
    assert reference_chunking(1000, 1) == (100, 10, 1000)
    assert reference_chunking(99, 1) == (100, 0, 0)


def test_counter_semantics():
    c = Counter()
    for lab in ["b", "a", "b", "Detected"]:
        c.increment_counter(lab)
    assert list(c.counter_dict) == ["b", "a", "Detected"] and c.counter_dict["b"] == 2
    assert c.calculate_efficiency() == 0.25
    d = Counter()
    d.merge_counters([c, c])
    assert d.counter_dict == {"b": 4, "a": 2, "Detected": 2}
    e = Counter.from_tallies(["x", "y", "z", "Detected"], [3, 0, 5, 1], [7, 2**62, 1, 9])
    assert list(e.counter_dict) == ["z", "x", "Detected"]  # first-seen order; zero counts absent
    assert Counter().calculate_efficiency() == 0


def test_beamline_sorting_and_packing():
    ap2 = CircularAperture("a2", 1.0, 0.01, d=0.01)
    ap1 = CircularAperture("a1", 0.5, 0.0, d=0.02)
    rect = RectangularAperture("r", 2.0, 0.1, x0=0.001, y0=-0.002, w=0.02, h=0.03)
    plates = FieldPlates("p", 1.5, 0.3, x0=0.0005, w=0.02)
    lens = ElectrostaticLens("L", 0.8, 0.6, d=0.035, V=30e3, stark=configs.stark_table(), field="hexapole",
                             integrator="rk4", dt=1e-5)
    bl = Beamline([ap2, rect, ap1, lens, plates], 9.81)
    assert [e.name for e in bl.elements] == ["a1", "L", "a2", "p", "r"]
    cb = compile_elements(bl.elements, bl.g)
    assert cb.labels == ["a1", "Lens entrance", "Inside lens", "a2", "p", "r", "Detected"]
    E = cb.elems
    assert E[0].kind == N.CIRC and E[0].p[0] == 0.01 and E[0].z1 == 0.5
    assert E[1].kind == N.LENS and E[1].force == N.FORCE_WSPLINE_HEX and E[1].integrator == N.INTEG_RK4
    cE = 3 * 30e3 / 0.0175 ** 3
    assert E[1].p[0] == 0.0175 and E[1].p[1] == 0.0175 * 0.0175 and E[1].p[2] == 1e-5 and E[1].p[3] == 0.6
    assert E[1].p[4] == pytest.approx(-2 * cE / lens.mass, rel=1e-15)
    assert E[1].p[5] == pytest.approx(cE / 1e4, rel=1e-15)
    assert E[1].z1 == 0.8 + 0.6
    assert E[3].kind == N.PLATES and (E[3].p[0], E[3].p[1]) == (0.0005 - 0.01, 0.0005 + 0.01)
    assert E[4].kind == N.RECT and (E[4].p[0], E[4].p[1], E[4].p[2], E[4].p[3]) == (
        0.001 - 0.01, 0.001 + 0.01, -0.002 - 0.015, -0.002 + 0.015)
    assert cb.phys.g == 9.81 and cb.phys.label_detected == 6 and cb.phys.n_labels == 7


def test_reference_lens_table_construction():
    """Without a_interp the lens builds its a(r) table as electrostatic_lens.py:193-206 does."""
    st = TwoLevelStark(sign=1.0)
    lens = ElectrostaticLens("L", 0.8, 0.6, state=st)
    r, a = lens.radial_table()
    d = 1.75 * 0.0254
    n = int(np.round(d / 2 / 1e-4))
    np.testing.assert_array_equal(r, np.linspace(0, d / 2 * 1.01, n))
    E = 2 * 27.6e3 / ((d / 2) ** 2) * r
    np.testing.assert_array_equal(a, -np.gradient(st.stark_potential(E / 100), 1e-4) / lens.mass)
    cb = compile_elements([lens], 9.80665)
    assert cb.elems[0].force == N.FORCE_RTABLE and cb.elems[0].n_steps == 600
    with pytest.raises(NotImplementedError):
        ElectrostaticLens("L", 0, 0.6).radial_table()
    with pytest.raises(ValueError):
        LinearTable(r, a)(r[-1] * 1.001)


def test_c_abi_library_exports_every_declared_symbol():
    """libctraj.so loads and exports every function include/ctraj.h declares (no GPU call made)."""
    hdr = (ROOT / "include" / "ctraj.h").read_text()
    declared = set(re.findall(r"^\s*(?:int32_t|const char \*)\s*(ctraj_\w+)\(", hdr, re.M))
    assert len(declared) >= 20
    lib = ctypes.CDLL(str(N.LIB_PATH))
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert {s for s, _, _ in N.SIGNATURES} == declared
    assert N.lib().ctraj_abi_version() == N.ABI_VERSION


_STRUCTS = {"Elem": "ctraj_elem", "Phys": "ctraj_phys", "SoA": "ctraj_soa", "Run": "ctraj_run",
            "Record": "ctraj_record", "BatchCfg": "ctraj_batch_cfg", "BatchTune": "ctraj_batch_tune",
            "Smooth": "ctraj_smooth"}


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Every ctypes mirror has the C struct's size and field offsets (gcc compiles a probe of the header)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ctraj.h"', 'int main(void) {']
    for py, c in _STRUCTS.items():
        lines.append(f'printf("{py} size %zu\\n", sizeof({c}));')
        for f, _ in getattr(N, py)._fields_:
            cf = {"inp": "in"}.get(f, f)  # Python keyword renamed in the mirror
            lines.append(f'printf("{py} {f} %zu\\n", offsetof({c}, {cf}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        py, f, v = line.split()
        S = getattr(N, py)
        got = ctypes.sizeof(S) if f == "size" else getattr(S, f).offset
        assert got == int(v), (py, f, got, v)


def test_library_is_newer_than_its_sources():
    """Guards against running a stale libctraj.so: rebuild (__graft_entry__.build()) after editing csrc/."""
    so = N.LIB_PATH
    srcs = list((ROOT / "centrex-molecule-trajectories_amd" / "csrc").glob("*")) + [ROOT / "include" / "ctraj.h"]
    newest = max(p.stat().st_mtime for p in srcs)
    assert so.exists() and so.stat().st_mtime >= newest - 1.0


def _two_level_hamiltonian(nu=6.67e9, mu_debye=4.23):
    from trajectories.stark import HBAR

    w = 2 * np.pi * nu
    c = mu_debye * DEBYE * 100.0 / HBAR  # rad/s per V/cm
    return np.diag([w / 2, -w / 2]), np.array([[0.0, c], [c, 0.0]])


@pytest.mark.parametrize("sign", [+1.0, -1.0])
def test_hamiltonian_stark_sweep_reproduces_the_two_level_closed_form(sign):
    """The reference's adiabatic sweep (stark_potential.py:9-90) on the two-level Hamiltonian
    H(E) = diag(+w/2, -w/2) + E mu (off-diagonal): the tracked state's energy minus its zero-field
    value is the closed form of TwoLevelStark (upper state: sign +1, lower: -1)."""
    from trajectories.stark import HamiltonianStark

    H0, HE = _two_level_hamiltonian()
    st = HamiltonianStark(H0, HE, [1.0, 0.0] if sign > 0 else [0.0, 1.0])
    Ezs = np.linspace(0.0, 70_000.0, 701)  # V/cm, as the reference sweeps
    V = st.stark_potential(Ezs)
    want = TwoLevelStark(sign=sign).energy(Ezs * 100.0)
    np.testing.assert_allclose(V - V[0], want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
    np.testing.assert_allclose(st.energy(Ezs * 100.0), V, rtol=0, atol=0)


def test_hamiltonian_stark_follows_diabatic_levels_through_crossings():
    """Uncoupled levels that cross: the overlap relabelling (reorder_evecs) keeps each level on its
    own straight line through every crossing, where sorted eigenvalues would switch lines."""
    from trajectories.stark import HBAR, HamiltonianStark

    H0 = np.diag([0.0, 10.0, 30.0])
    HE = np.diag([1.0, 0.5, -0.2])
    Ezs = np.linspace(0.3, 99.7, 200)
    for i in range(3):
        st = HamiltonianStark(H0, HE, np.eye(3)[i])
        np.testing.assert_allclose(st.stark_potential(Ezs) / HBAR, H0[i, i] + HE[i, i] * Ezs, rtol=1e-12, atol=1e-12)
    sorted_low = np.linalg.eigvalsh(H0[None] + Ezs[:, None, None] * HE[None])[:, 0]
    assert not np.allclose(sorted_low, H0[0, 0] + HE[0, 0] * Ezs)  # the lowest level is not one line


def test_lens_tables_from_the_hamiltonian_sweep_equal_the_two_level_ones():
    """A lens whose state is the two-level Hamiltonian sweep builds the a(r) table (electrostatic_lens.py:
    193-206) and the W' spline of the TwoLevelStark closed form, to the eigensolver's rounding."""
    from trajectories.stark import HamiltonianStark

    H0, HE = _two_level_hamiltonian()
    ham = HamiltonianStark(H0, HE, [1.0, 0.0])
    r1, a1 = ElectrostaticLens("L", 0.8, 0.6, state=ham).radial_table()
    r2, a2 = ElectrostaticLens("L", 0.8, 0.6, state=TwoLevelStark(sign=1.0)).radial_table()
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_allclose(a1, a2, rtol=1e-7, atol=1e-9 * np.abs(a2).max())
    s1 = WSpline.from_stark(ham, 100e3, 201)
    s2 = WSpline.from_stark(TwoLevelStark(sign=1.0), 100e3, 201)
    np.testing.assert_allclose(s1.coef[:, :3], s2.coef[:, :3], rtol=1e-6, atol=1e-9 * np.abs(s2.coef).max())
