"""Per-molecule numpy restatement of the reference's propagation.  TEST INFRASTRUCTURE ONLY.

It keeps the reference's structure and numpy arithmetic: one molecule at a time, 3-vectors, a list of
trajectory rows, `delta_t**2` on a numpy scalar (libm pow), np.sqrt(np.sum(x[:2]**2)), np.interp for
the lens table -- so it reproduces the reference's floating-point results, not just its algorithm.
It reads element fields from the product's element objects (name, z0, z1, d, x1.., lens fields).

Used by tests (cross-check of the C oracle and of the packed element table) and by bench.py's
cpu_baseline leg (``run_simulation_parallel``: the reference's joblib chunking over host cores).

Follows:
  molecule.py:26-68 (x(dt), v(dt), update_trajectory), beamline.py:20-35 (walk),
  apertures.py:92-115 / 165-189 / 227-270 (elements), electrostatic_lens.py:48-118, 215-228 (lens),
  trajectory_simulator.py:48-97 (chunking, per-chunk draws, counters).
The configuration-mode lens (hexapole W(|E|) spline, classical RK4 at fixed dt with a final partial
step) is restated from its definition with plain numpy arithmetic (no fused multiply-adds).
"""
from __future__ import annotations

import numpy as np

LENS_ENTRANCE, INSIDE_LENS, DETECTED = "Lens entrance", "Inside lens", "Detected"


class Mol:
    def __init__(self, x0, v0, a0, t0):
        self.X = [np.array(x0, dtype=np.float64)]
        self.Vv = [np.array(v0, dtype=np.float64)]
        self.A = [np.array(a0, dtype=np.float64)]
        self.T = [np.float64(t0)]
        self.alive = True
        self.hit = None

    def x(self, dt=None):
        if not dt:
            return self.X[-1]
        return self.x() + self.v() * dt + self.a() * dt ** 2 / 2

    def v(self, dt=None):
        if not dt:
            return self.Vv[-1]
        return self.v() + self.a() * dt

    def a(self):
        return self.A[-1]

    def t(self):
        return self.T[-1]

    def row(self, x, v, a, t):
        self.X.append(np.array(x, dtype=np.float64))
        self.Vv.append(np.array(v, dtype=np.float64))
        self.A.append(np.array(a, dtype=np.float64))
        self.T.append(t)

    def update(self, dt, g):
        x = self.x(dt)
        v = self.v(dt)
        t = self.t() + dt
        self.row(x, v, np.array((0, -g, 0)), t)

    def kill(self, name):
        self.alive = False
        self.hit = name


def _circ(el, m, g):
    for z in (el.z0, el.z1):
        m.update((z - m.x()[2]) / m.v()[2], g)
        if np.sqrt(np.sum(m.x()[:2] ** 2)) > el.d / 2:
            m.kill(el.name)
            return


def _rect(el, m, g):
    for z in (el.z0, el.z1):
        m.update((z - m.x()[2]) / m.v()[2], g)
        if not ((el.x1 < m.x()[0] < el.x2) and (el.y1 < m.x()[1] < el.y2)):
            m.kill(el.name)
            return


def _plates(el, m, g):
    m.update((el.z0 - m.x()[2]) / m.v()[2], g)
    if not (el.x1 < m.x()[0] < el.x2):
        m.kill(el.name)
        return
    dt = (el.z1 - m.x()[2]) / m.v()[2]
    x = m.x(dt)
    if not (el.x1 < x[0] < el.x2):
        if m.v()[0] < 0:
            dt = (el.x1 - m.x()[0]) / m.v()[0]
        elif m.v()[0] > 0:
            dt = (el.x2 - m.x()[0]) / m.v()[0]
        m.update(dt, g)
        m.kill(el.name)
        return
    m.update(dt, g)


class LensForce:
    """Acceleration inside a lens: the reference's a_r(r) table (np.interp), or -grad W(|E|)/m."""

    def __init__(self, el, g):
        self.g = g
        self.el = el
        if el.stark is None:
            self.rr, self.aa = el.radial_table()
            self.kind = "rtable"
        else:
            spl = el.stark if not isinstance(el.stark, (list, tuple)) else el.stark
            self.splines = spl if isinstance(spl, (list, tuple)) else [spl]
            self.cE = el.field_coefficient()
            self.kind = el.field

    def __call__(self, x, state=0):
        g = self.g
        a = np.zeros((3,))
        if self.kind == "rtable":
            r = np.sqrt(np.sum(x[:2] ** 2))
            if r < self.rr[0] or r > self.rr[-1]:
                raise ValueError(f"A value ({r}) in x_new is outside the interpolation range.")
            a_r = np.interp(r, self.rr, self.aa)
            if r != 0:
                a[0] = a_r * x[0] / r
                a[1] = a_r * x[1] / r
                a[2] = 0
        else:
            spl = self.splines[state]
            r2 = x[0] * x[0] + x[1] * x[1]
            if self.kind == "hexapole":
                u = min(r2 * (self.cE / spl.dE), np.nextafter(float(spl.n_seg), 0.0))
                k = int(u)
                s = u - k
                A, B, C = spl.coef[k, 0], spl.coef[k, 1], spl.coef[k, 2]
                f = ((A * s + B) * s + C) * (-2.0 * self.cE / self.el.mass)
                a[0] = f * x[0]
                a[1] = f * x[1]
            else:
                r = np.sqrt(r2)
                if r != 0:
                    u = min(r * (self.cE / spl.dE), np.nextafter(float(spl.n_seg), 0.0))
                    k = int(u)
                    s = u - k
                    A, B, C = spl.coef[k, 0], spl.coef[k, 1], spl.coef[k, 2]
                    f = ((A * s + B) * s + C) * (-self.el.mass ** -1 * self.cE) / r
                    a[0] = f * x[0]
                    a[1] = f * x[1]
        a[1] -= g
        return a


def _lens(el, m, g, force, state):
    m.update((el.z0 - m.x()[2]) / m.v()[2], g)
    if np.sqrt(np.sum(m.x()[:2] ** 2)) > el.d / 2:
        m.kill(LENS_ENTRANCE)
        return
    if el.integrator == "ref":
        N = int(np.rint(el.L / el.dz))
        dt = el.dz / m.v()[2]
        for _ in range(N):
            x = m.x()
            k1 = m.v()
            l1 = force(x, state)
            k2 = k1 + dt * l1 / 2
            l2 = force(x + dt * k1, state)
            k3 = k1 + dt * l2 / 2
            l3 = force(x + dt * k2 / 2, state)
            k4 = k1 + dt * l3
            l4 = force(x + dt * k3, state)
            m.row(x + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6, k1 + dt * (l1 + 2 * l2 + 2 * l3 + l4) / 6, l1,
                  m.t() + dt)
            if np.sqrt(np.sum(m.x()[:2] ** 2)) > el.d / 2:
                m.kill(INSIDE_LENS)
                return
        m.update((el.z1 - m.x()[2]) / m.v()[2], g)
        return
    # configuration mode: classical RK4 on (x, y, vx, vy) at fixed dt, then a partial step to z1
    h = el.dt
    vz = m.v()[2]
    q = np.floor(el.L / (vz * h))
    nfull = int(q) if q >= 0 else 0
    x = np.array(m.x()[:2])
    v = np.array(m.v()[:2])
    z, t = m.x()[2], m.t()
    R2 = (el.d / 2) * (el.d / 2)

    def f(p):
        return force(np.array((p[0], p[1], 0.0)), state)[:2]

    for i in range(nfull + 1):
        hs = h if i < nfull else el.L / vz - nfull * h
        a1 = f(x)
        a2 = f(x + hs / 2 * v)
        v2 = v + hs / 2 * a1
        a3 = f(x + hs / 2 * v2)
        v3 = v + hs / 2 * a2
        a4 = f(x + hs * v3)
        v4 = v + hs * a3
        x = x + hs / 6 * (v + 2 * v2 + 2 * v3 + v4)
        v = v + hs / 6 * (a1 + 2 * a2 + 2 * a3 + a4)
        z = z + vz * h if i < nfull else el.z1
        t = t + hs
        last = i == nfull
        inside = x[0] * x[0] + x[1] * x[1] > R2
        acc = np.array((0, -g, 0)) if (last and not inside) else force(np.array((x[0], x[1], 0.0)), state)
        m.row((x[0], x[1], z), (v[0], v[1], vz), acc, t)
        if inside:
            m.kill(INSIDE_LENS)
            return


def propagate_one(elements, x0, v0, g, forces=None, state=0, a0=None, t0=0.0):
    """One molecule through the z0-sorted elements; returns the Mol (rows, alive, hit)."""
    m = Mol(x0, v0, np.array((0, -g, 0)) if a0 is None else a0, t0)
    for el in elements:
        kind = type(el).__name__
        if kind == "CircularAperture":
            _circ(el, m, g)
        elif kind == "RectangularAperture":
            _rect(el, m, g)
        elif kind == "FieldPlates":
            _plates(el, m, g)
        elif kind == "ElectrostaticLens":
            _lens(el, m, g, forces[id(el)], state)
        else:
            raise TypeError(kind)
        if not m.alive:
            break
    if m.alive:
        m.hit = DETECTED
    return m


def make_forces(elements, g):
    return {id(el): LensForce(el, g) for el in elements if type(el).__name__ == "ElectrostaticLens"}


def propagate_block(elements, g, X, V, states=None):
    """Molecules X[:, i], V[:, i] one by one; returns (labels list, final rows (n, 7))."""
    forces = make_forces(elements, g)
    hits = []
    fin = np.empty((X.shape[1], 7))
    for i in range(X.shape[1]):
        m = propagate_one(elements, X[:, i], V[:, i], g, forces, 0 if states is None else int(states[i]))
        hits.append(m.hit)
        fin[i, :3] = m.x()
        fin[i, 3:6] = m.v()
        fin[i, 6] = m.t()
    return hits, fin


def _chunk(elements, g, X, V, aoi):
    counter = {}
    kept = []
    hits, fin = propagate_block(elements, g, X, V)
    for i, h in enumerate(hits):
        counter[h] = counter.get(h, 0) + 1
        if h in aoi:
            kept.append(fin[i])
    return counter, kept


def run_simulation_parallel(elements, g, X, V, apertures_of_interest=(), n_jobs=1):
    """The reference's run_simulation parallel section (trajectory_simulator.py:48-97): N_loops = 100*n_jobs
    chunks of the given molecules run by joblib workers; per-chunk counters merged in chunk order."""
    from joblib import Parallel, delayed

    n = X.shape[1]
    loops = 100 * n_jobs
    N = n // loops
    parts = Parallel(n_jobs=n_jobs)(
        delayed(_chunk)(elements, g, X[:, c * N:(c + 1) * N], V[:, c * N:(c + 1) * N], tuple(apertures_of_interest))
        for c in range(loops))
    counter = {}
    kept = []
    for c, k in parts:
        for key, val in c.items():
            counter[key] = counter.get(key, 0) + val
        kept.extend(k)
    return counter, kept
