"""The benchmark beamlines, with every constant the workload descriptions leave open pinned here.

Source (all workloads): x, y ~ N(0, 1.5 mm), z = 0; vx, vy ~ N(0, 1.5 m/s); vz ~ N(184, 16) m/s,
drawn in the reference's order (per chunk: vx, vy, vz, then x, y).  Gravity 9.81 m/s^2 along -Y.
Aperture thickness L = 0 (thin).  Lens: hexapole |E| = 3 V0 r^2 / R^3, R = 1.75 cm, molecular mass
= the reference default (204.38 + 19.00) * 1.67e-27 kg, W(|E|) two-level with nu = 6.67 GHz,
mu = 4.23 D, tabulated at 1001 knots over 0..100 kV/cm and splined (not-a-knot), classical RK4 at
dt = 10 us with a final partial step.  Aperture sizes below are radii.
"""
from __future__ import annotations

import numpy as np

from .beamline import Beamline
from .beamline_elements import CircularAperture, ElectrostaticLens, FieldPlates, RectangularAperture
from .distributions import CeNTREXVelocityDistribution, GaussianPositionDistribution
from .stark import TwoLevelStark, WSpline

G = 9.81
SIGMA_X = 1.5e-3
SIGMA_V = 1.5
VZ_MEAN, VZ_SIGMA = 184.0, 16.0
LENS_R = 1.75e-2
LENS_L = 0.6
LENS_DT = 10e-6
NU = 6.67e9
MU_D = 4.23

_TABLES: dict = {}


def source():
    """(xdist, vdist) of the benchmark source."""
    return (GaussianPositionDistribution(SIGMA_X, SIGMA_X, 0.0),
            CeNTREXVelocityDistribution(0.0, SIGMA_V, 0.0, SIGMA_V, VZ_MEAN, VZ_SIGMA))


def draw_source(n: int, seed: int, n_jobs: int = 1):
    """Host draw of n molecules with np.random.RandomState(seed) in the reference's order.
    Returns X (3, n), V (3, n).  n must be a multiple of 100*n_jobs (the reference's chunking)."""
    xdist, vdist = source()
    rs = np.random.RandomState(seed)
    loops = 100 * n_jobs
    N = n // loops
    if N * loops != n:
        raise ValueError("n must be a multiple of 100*n_jobs")
    X = np.empty((3, n))
    V = np.empty((3, n))
    for c in range(loops):
        s = slice(c * N, (c + 1) * N)
        V[0, s] = vdist.vx + vdist.sigmax * rs.standard_normal(N)
        V[1, s] = vdist.vy + vdist.sigmay * rs.standard_normal(N)
        V[2, s] = vdist.vz + vdist.sigmaz * rs.standard_normal(N)
        X[0, s] = 0.0 + xdist.sigmax * rs.standard_normal(N)
        X[1, s] = 0.0 + xdist.sigmay * rs.standard_normal(N)
        X[2, s] = xdist.z
    return X, V


def stark_table(sign: float = -1.0, dipole_scale: float = 1.0) -> WSpline:
    """W(|E|) spline of the two-level state (cached)."""
    key = (float(sign), float(dipole_scale))
    if key not in _TABLES:
        _TABLES[key] = WSpline.from_stark(TwoLevelStark(NU, MU_D, sign, dipole_scale), 100e3, 1001)
    return _TABLES[key]


def config0() -> Beamline:
    """Ballistic beamline: apertures r = 6, 4, 2.5 mm at z = 0.3, 0.8, 1.2 m; detector r = 5 mm at 2.0 m."""
    return Beamline([
        CircularAperture("Aperture 1", 0.3, 0.0, d=2 * 6e-3),
        CircularAperture("Aperture 2", 0.8, 0.0, d=2 * 4e-3),
        CircularAperture("Aperture 3", 1.2, 0.0, d=2 * 2.5e-3),
        CircularAperture("Detector", 2.0, 0.0, d=2 * 5e-3),
    ], G)


def lens(V0: float = 30e3, z0: float = 0.8, stark=None, integrator: str = "rk4", dt: float = LENS_DT):
    return ElectrostaticLens("ES lens", z0, LENS_L, d=2 * LENS_R, V=V0, stark=stark or stark_table(),
                             field="hexapole", integrator=integrator, dt=dt)


def config1(V0: float = 30e3, lens_z0: float = 0.8, r1: float = 6e-3, r2: float = 3e-3, r3: float = 4e-3,
            stark=None) -> Beamline:
    """ES-lens beamline: aperture r1 at 0.3 m; hexapole lens at z0 (L = 0.6 m); aperture r2 at 1.6 m;
    detector r3 at 2.5 m."""
    return Beamline([
        CircularAperture("Aperture 1", 0.3, 0.0, d=2 * r1),
        lens(V0, lens_z0, stark),
        CircularAperture("Aperture 2", 1.6, 0.0, d=2 * r2),
        CircularAperture("Detector", 2.5, 0.0, d=2 * r3),
    ], G)


def config4_tables():
    """Four states: the two-level form with sign -1 and +1, and sign -1 with the dipole scaled by 0.5, 0.8."""
    return [stark_table(-1.0, 1.0), stark_table(+1.0, 1.0), stark_table(-1.0, 0.5), stark_table(-1.0, 0.8)]


def config4() -> Beamline:
    return config1(stark=config4_tables())


SOURCE_MEAN = (0.0, 0.0, 0.0, 0.0, 0.0, VZ_MEAN)
SOURCE_SIGMA = (SIGMA_X, SIGMA_X, 0.0, SIGMA_V, SIGMA_V, VZ_SIGMA)


def config2(n_cfg: int = 256, seed: int = 2):
    """The batched scan: (beamlines, counts, device source seeds).  Configuration c is the ES-lens
    beamline with its own V0, lens z0 and final (detector) radius; its molecules come from the device
    source with seed (seed << 32) + c."""
    V0, z0, r, n = config2_params(n_cfg, seed)
    bls = [config1(V0=float(V0[c]), lens_z0=float(z0[c]), r3=float(r[c])) for c in range(n_cfg)]
    seeds = (np.uint64(seed) << np.uint64(32)) + np.arange(n_cfg, dtype=np.uint64)
    return bls, n, seeds


def config2_params(n_cfg: int = 256, seed: int = 2):
    """Per-configuration (V0, lens z0, final aperture radius, molecule count), drawn from
    RandomState(seed) in that order: V0 ~ U[10, 40] kV, z0 ~ U[0.7, 0.9] m, r ~ U[2, 5] mm,
    n = round(10**U[3, 6])."""
    rs = np.random.RandomState(seed)
    V0 = rs.uniform(10e3, 40e3, n_cfg)
    z0 = rs.uniform(0.7, 0.9, n_cfg)
    r = rs.uniform(2e-3, 5e-3, n_cfg)
    n = np.round(10.0 ** rs.uniform(3.0, 6.0, n_cfg)).astype(np.int64)
    return V0, z0, r, n


# Differentiable objective: config-1 geometry, theta = (r1, V0, lens z0, r2, r3, mu_vz) in beamline
# order, sigmoid width 50 um, lens checkpoints every 8 steps (each checkpoint segment is then one
# recompute window: every step is recomputed once; 64 steps took 7.5 ms against 6.3 ms).
CONFIG3_SOFTNESS = 50e-6
CONFIG3_CKPT = 64


def config3_source(n: int, seed: int = 3):
    """(x, y, vx, vy, eps) of the differentiable workload: the benchmark source with vz
    reparameterised as mu + 16 eps (eps standard normal, drawn where the source draws vz)."""
    X, V = draw_source(n, seed)
    rs = np.random.RandomState(seed)
    loops = 100
    N = n // loops
    eps = np.empty(n)
    for c in range(loops):
        rs.standard_normal(N)
        rs.standard_normal(N)
        eps[c * N:(c + 1) * N] = rs.standard_normal(N)
        rs.standard_normal(N)
        rs.standard_normal(N)
    return X[0], X[1], V[0], V[1], eps


def config3_theta() -> np.ndarray:
    return np.array([6e-3, 30e3, 0.8, 3e-3, 4e-3, VZ_MEAN])


def reference_example(sign: float = +1.0) -> Beamline:
    """The beamline of the reference's examples/lens_simulation_beamline.py:18-72 with its lens in the
    reference's own model (integrator 'ref', linear field, a(r) table built as
    electrostatic_lens.py:193-209 does), here from the two-level Stark form of the given sign."""
    m = 0.0254
    fourK = CircularAperture(z0=1.7 * m, L=0.25 * m, d=1 * m, name="4K shield")
    fortyK = CircularAperture(z0=fourK.z1 + 1.25 * m, L=0.25 * m, d=1 * m, name="40K shield")
    bb_exit = CircularAperture(z0=fortyK.z1 + 2.5 * m, L=0.75 * m, d=4 * m, name="BB exit")
    lens = ElectrostaticLens(z0=bb_exit.z1 + 33 * m, L=0.6, name="ES lens", state=TwoLevelStark(sign=sign))
    plates = FieldPlates(z0=2.43, L=3.0, w=0.02, name="Field plates")
    dr = RectangularAperture(z0=plates.z1 + 39.9 * m, L=0.25 * m, name="DR aperture", w=0.018, h=0.03)
    return Beamline([fourK, fortyK, bb_exit, lens, plates, dr])
