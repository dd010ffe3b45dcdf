"""Stark potential W(|E|) tables for the electrostatic lens.

The reference builds its lens force from ``stark_potential(state, E)`` -- a diagonalisation of the
TlF Hamiltonian through ``centrex_TlF`` (stark_potential.py:9-67) -- sampled at
``electrostatic_lens.py:194-206``.  That package is not part of this engine.  Two table forms are
provided instead:

* :class:`TwoLevelStark` -- the closed form W(E) = sign * (sqrt((h nu/2)^2 + (a mu E)^2) - h nu/2).
  ``sign=-1`` is the literal form of the benchmark configuration (a high-field seeker).  It also
  works as the lens ``state`` for the reference's own a(r)-table construction.
* :class:`HamiltonianStark` -- the reference's own construction (stark_potential.py:9-90) for a
  Hamiltonian H(E) = H0 + E HE given as matrices: a sweep of diagonalisations over |E| with the
  eigenvectors relabelled adiabatically (reorder_evecs) and the state picked by its overlap with a
  given state vector.  With the TlF Hamiltonian (which needs centrex_TlF to build) this is the
  reference's table; with any other Hamiltonian it is the same algorithm.
* :class:`WSpline` -- a not-a-knot cubic spline of W on a uniform |E| grid, stored per segment as
  the coefficients of W'(E) = (A s + B) s + C in the segment coordinate s = E/dE - k.  This is the
  table the device stages in LDS.
"""
from __future__ import annotations

import numpy as np

H_PLANCK = 6.62607015e-34  # J s
HBAR = H_PLANCK / (2.0 * np.pi)  # J s
C_LIGHT = 299792458.0  # m/s
DEBYE = 1e-21 / C_LIGHT  # C m


class TwoLevelStark:
    """W(E) = sign * (sqrt((h nu / 2)^2 + (dipole_scale * mu * E)^2) - h nu / 2), E in V/m, W in J."""

    def __init__(self, nu: float = 6.67e9, mu_debye: float = 4.23, sign: float = -1.0, dipole_scale: float = 1.0):
        self.nu = float(nu)
        self.mu_debye = float(mu_debye)
        self.sign = float(sign)
        self.dipole_scale = float(dipole_scale)
        self.half_hnu = H_PLANCK * self.nu / 2.0
        self.mu = self.dipole_scale * self.mu_debye * DEBYE

    def energy(self, E):
        E = np.asarray(E, dtype=np.float64)
        return self.sign * (np.sqrt(self.half_hnu ** 2 + (self.mu * E) ** 2) - self.half_hnu)

    def denergy(self, E):
        """dW/dE (analytic)."""
        E = np.asarray(E, dtype=np.float64)
        return self.sign * self.mu ** 2 * E / np.sqrt(self.half_hnu ** 2 + (self.mu * E) ** 2)

    # the reference calls stark_potential(state, E_values / 100) with E in V/cm (electrostatic_lens.py:203)
    def stark_potential(self, E_V_per_cm):
        return self.energy(np.asarray(E_V_per_cm, dtype=np.float64) * 100.0)

    def __repr__(self):
        return (f"TwoLevelStark(nu={self.nu!r}, mu_debye={self.mu_debye!r}, sign={self.sign!r}, "
                f"dipole_scale={self.dipole_scale!r})")


def reorder_evecs(V_in: np.ndarray, E_in: np.ndarray, V_ref: np.ndarray):
    """Relabel one diagonalisation adiabatically (stark_potential.py:70-90): eigenvector i of V_in
    (columns) is matched to the column of V_ref it overlaps most, and the eigenpairs are put in
    the order of their matches, so that index k keeps labelling the same state along a sweep."""
    overlap = np.abs(V_in.conj().T @ V_ref)
    index = np.argsort(np.argmax(overlap, axis=1))
    return E_in[index], V_in[:, index]


def closest_vector_index(state_vector: np.ndarray, V: np.ndarray) -> int:
    """The column of V (eigenvectors) with the largest overlap |<v|state>| with the state vector."""
    return int(np.argmax(np.abs(V.conj().T @ np.asarray(state_vector))))


class HamiltonianStark:
    """Stark potential of one eigenstate of H(E) = H0 + E HE, E along z in V/cm, H in rad/s, by the
    reference's sweep (stark_potential.py:9-67): the eigenvectors at ``E_ref`` (100 V/cm in the
    reference) are the initial labels and the state is the one closest to ``state_vector`` there;
    each field of a sweep is diagonalised (numpy's eigh, as the reference) and relabelled against the
    previous one; the potential is the state's eigenvalue times hbar (J).

    ``H0`` may include a small magnetic term to split degeneracies, as the reference's H(E, B) does.
    ``stark_potential(E_V_per_cm)`` takes the fields in sweep order (the labels follow the path);
    ``energy(E_V_per_m)`` is the same for :meth:`WSpline.from_stark`."""

    def __init__(self, H0, HE, state_vector, E_ref: float = 100.0):
        self.H0 = np.asarray(H0)
        self.HE = np.asarray(HE)
        if self.H0.ndim != 2 or self.H0.shape[0] != self.H0.shape[1] or self.HE.shape != self.H0.shape:
            raise ValueError("H0 and HE must be square matrices of one shape")
        self.E_ref = float(E_ref)
        self.state_vector = np.asarray(state_vector)
        _, self.V_ref0 = np.linalg.eigh(self.H0 + self.E_ref * self.HE)
        self.index = closest_vector_index(self.state_vector, self.V_ref0)

    def stark_potential(self, E_V_per_cm):
        Ezs = np.atleast_1d(np.asarray(E_V_per_cm, dtype=np.float64))
        D, V = np.linalg.eigh(self.H0[None, :, :] + Ezs[:, None, None] * self.HE[None, :, :])
        out = np.empty(Ezs.size)
        V_ref = self.V_ref0
        for i in range(Ezs.size):
            energies, V_ref = reorder_evecs(V[i], D[i], V_ref)
            out[i] = energies[self.index]
        out *= HBAR
        return out.reshape(np.shape(E_V_per_cm))

    def energy(self, E):
        return self.stark_potential(np.asarray(E, dtype=np.float64) / 100.0)

    def __repr__(self):
        return f"HamiltonianStark(dim={self.H0.shape[0]}, state_index={self.index}, E_ref={self.E_ref!r})"


def not_a_knot_second_derivatives(y: np.ndarray, h: float) -> np.ndarray:
    """Second derivatives M_i of the not-a-knot cubic spline through y on a uniform grid of step h.

    With equal spacing the not-a-knot conditions (third derivative continuous at the second and
    second-to-last knots) give M_1 and M_{n-1} as plain second differences; the interior rows
    M_{i-1} + 4 M_i + M_{i+1} = 6 (y_{i+1} - 2 y_i + y_{i-1}) / h^2 are then tridiagonal.
    """
    y = np.asarray(y, dtype=np.float64)
    n = y.size - 1  # segments
    if n < 4:
        raise ValueError("need at least 5 knots")
    d2 = (y[2:] - 2.0 * y[1:-1] + y[:-2]) / (h * h)  # second differences at knots 1..n-1
    M = np.empty(n + 1)
    M[1] = d2[0]
    M[n - 1] = d2[-1]
    # rows i = 2 .. n-2
    m = n - 3
    if m > 0:
        rhs = 6.0 * d2[1:-1].copy()
        rhs[0] -= M[1]
        rhs[-1] -= M[n - 1]
        # Thomas algorithm for diag 4, off-diagonals 1
        c = np.empty(m)
        d = np.empty(m)
        c[0] = 1.0 / 4.0
        d[0] = rhs[0] / 4.0
        for i in range(1, m):
            den = 4.0 - c[i - 1]
            c[i] = 1.0 / den
            d[i] = (rhs[i] - d[i - 1]) / den
        sol = np.empty(m)
        sol[-1] = d[-1]
        for i in range(m - 2, -1, -1):
            sol[i] = d[i] - c[i] * sol[i + 1]
        M[2:n - 1] = sol
    M[0] = 2.0 * M[1] - M[2]
    M[n] = 2.0 * M[n - 1] - M[n - 2]
    return M


class WSpline:
    """Cubic spline of W(E) on the uniform grid E_k = k dE, k = 0..n_seg, stored as W' coefficients."""

    def __init__(self, E_knots, W_values):
        E = np.asarray(E_knots, dtype=np.float64)
        W = np.asarray(W_values, dtype=np.float64)
        if E.ndim != 1 or E.shape != W.shape or E.size < 5:
            raise ValueError("E_knots and W_values must be 1-D of equal length >= 5")
        if E[0] != 0.0:
            raise ValueError("the |E| grid must start at 0")
        h = (E[-1] - E[0]) / (E.size - 1)
        if not np.allclose(np.diff(E), h, rtol=1e-9, atol=0.0):
            raise ValueError("the |E| grid must be uniform")
        self.E = E
        self.W = W
        self.dE = float(h)
        self.n_seg = E.size - 1
        M = not_a_knot_second_derivatives(W, h)
        self.M = M
        # S(t) = y_i + b t + c t^2 + d t^3 on [E_i, E_i+1], t = E - E_i
        self.c2 = M[:-1] / 2.0
        self.c3 = (M[1:] - M[:-1]) / (6.0 * h)
        self.c1 = (W[1:] - W[:-1]) / h - h * (2.0 * M[:-1] + M[1:]) / 6.0
        self.c0 = W[:-1].copy()
        coef = np.zeros((self.n_seg, 4))
        coef[:, 0] = 3.0 * self.c3 * h * h  # A
        coef[:, 1] = 2.0 * self.c2 * h  # B
        coef[:, 2] = self.c1  # C
        self.coef = coef

    @classmethod
    def from_stark(cls, stark, E_max_V_per_cm: float = 100e3, n_knots: int = 1001) -> "WSpline":
        """Tabulate ``stark.energy`` at n_knots points over 0..E_max (V/cm) and spline it."""
        E = np.linspace(0.0, E_max_V_per_cm, n_knots) * 100.0  # V/m
        out = cls(E, stark.energy(E))
        out.stark = stark
        return out

    # host evaluation (tests, plotting) -- the same segment rule as the device
    def _seg(self, E):
        u = np.minimum(np.asarray(E, dtype=np.float64) / self.dE, np.nextafter(float(self.n_seg), 0.0))
        k = u.astype(np.int64)
        return u, k, u - k

    def dW(self, E):
        u, k, s = self._seg(E)
        A, B, Cc = self.coef[k, 0], self.coef[k, 1], self.coef[k, 2]
        return (A * s + B) * s + Cc

    def W_at(self, E):
        u, k, s = self._seg(E)
        t = s * self.dE
        return self.c0[k] + t * (self.c1[k] + t * (self.c2[k] + t * self.c3[k]))
