"""Initial position / velocity distributions (distributions.py:12-185 of the reference).

``draw(n)`` is the reference's host draw (scipy.stats on numpy's global RNG), so a seeded run
replays the reference's inputs exactly.  ``draw_device(n, seed)`` (Gaussian families) draws on the
GPU with Philox4x32-10 + Box-Muller into device buffers, for runs too large to ship from the host.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass
from pathlib import Path

import numpy as np
from scipy.stats import norm, uniform


class Distribution(ABC):
    @abstractmethod
    def draw(self, n: int) -> np.ndarray:
        """n samples as a (3, n) float64 array."""

    def save_to_hdf(self, filepath: Path, run_name: str):
        from .hdf import save_distribution

        save_distribution(self, filepath, run_name)

    # device forms: (mean[3], sigma[3]) of independent Gaussians, or None if not Gaussian
    def gaussian_params(self):
        return None

    # (radius, z) of a uniform disc at fixed z, or None if not such a disc
    def disc_params(self):
        return None

    def draw_device(self, n: int, seed: int, offset: int = 0, ctx=None, out=None):
        """Draw n samples on the device from the counter-based streams of key ``seed`` (molecule
        index + ``offset`` as the counter, so any chunking of a run draws the same molecules);
        returns three DeviceArrays (rows of draw(n); ``out``, three arrays of at least n, to fill
        existing ones).  The values follow the distribution, not the host draw's numpy sequence."""
        from .device import get_context
        from .engine import draw_disc, draw_gaussian

        p = self.gaussian_params()
        if p is not None:
            mean, sigma = p
            return draw_gaussian(ctx or get_context(), n, mean, sigma, seed, offset, out=out)
        q = self.disc_params()
        if q is not None:
            radius, z = q
            return draw_disc(ctx or get_context(), n, radius, z, seed, offset, out=out)
        raise NotImplementedError(f"{type(self).__name__} has no device sampler")


@dataclass
class GaussianDistribution(Distribution):
    mean: float
    sigma: float

    def __post_init__(self):
        self.dist = norm

    def draw(self, n: int) -> np.ndarray:
        return self.dist.rvs(loc=self.mean, scale=self.sigma, size=n)

    def plot(self):
        import matplotlib.pyplot as plt

        fig, ax = plt.subplots()
        x = self.mean + np.linspace(-5, 5, 1000) * self.sigma
        ax.plot(x, self.dist.pdf(x, loc=self.mean, scale=self.sigma))

    def save_to_hdf(self, filepath: Path, run_name: str, group_name: str = ""):
        raise NotImplementedError("Saving GaussianDistribution to hdf is not implemented. Save child class instead.")


@dataclass
class CeNTREXVelocityDistribution(Distribution):
    """Independent Gaussians along X, Y, Z; draw order vx, vy, vz (distributions.py:69-76)."""

    vx: float = 0.0
    sigmax: float = 39.5
    vy: float = 0.0
    sigmay: float = 39.5
    vz: float = 184.0
    sigmaz: float = 16.0

    def draw(self, n: int) -> np.ndarray:
        return np.array((GaussianDistribution(self.vx, self.sigmax).draw(n),
                         GaussianDistribution(self.vy, self.sigmay).draw(n),
                         GaussianDistribution(self.vz, self.sigmaz).draw(n)))

    def gaussian_params(self):
        return (np.array([self.vx, self.vy, self.vz], float), np.array([self.sigmax, self.sigmay, self.sigmaz], float))


@dataclass
class CeNTREXPositionDistribution(Distribution):
    """Uniform disc of diameter d at z (distributions.py:112-119): theta then sqrt(u)."""

    d: float = 0.02
    z: float = 0.25 * 0.0254

    def draw(self, n: int) -> np.ndarray:
        theta = uniform.rvs(size=n) * 2 * np.pi
        r = np.sqrt(uniform.rvs(size=n)) * self.d / 2
        return np.array((r * np.cos(theta), r * np.sin(theta), np.full(n, self.z)))

    def disc_params(self):
        return (self.d / 2, float(self.z))


@dataclass
class GaussianPositionDistribution(Distribution):
    """Gaussian x, y at fixed z; draw order x, y (distributions.py:155-162)."""

    sigmax: float = 0.25 * 25.4 / 5 * 3.8e-3
    sigmay: float = 0.25 * 25.4 / 5 * 3.8e-3
    z: float = 0.25 * 0.0254

    def draw(self, n: int) -> np.ndarray:
        x = GaussianDistribution(0, self.sigmax).draw(n)
        y = GaussianDistribution(0, self.sigmay).draw(n)
        return np.vstack((x, y, np.full(n, self.z)))

    def gaussian_params(self):
        return (np.array([0.0, 0.0, self.z], float), np.array([self.sigmax, self.sigmay, 0.0], float))
