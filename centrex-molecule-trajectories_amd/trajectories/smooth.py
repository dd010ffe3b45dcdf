"""Differentiable detected flux: the sigmoid-smoothed objective and its hand-written adjoint kernel.

The reference has no gradient; this is the differentiable form of its Monte Carlo flux estimate
(``TrajectorySimulator.run_simulation`` counting "Detected", trajectory_simulator.py:34-103):

    J(theta) = (1/n) sum_i prod_{aperture faces f} sigmoid((r_f - rho_i,f) / softness)

with every molecule walked through the whole beamline (no molecule is stopped), vz_i = mu + sigma_vz
eps_i, and theta = (each circular aperture's radius, each lens's (V, z0), mu) in beamline order.
``ctraj_smooth_objective`` computes J and dJ/dtheta for one or many parameter sets in one launch;
``smooth_flux`` wraps it in a ``torch.autograd.Function``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .device import DeviceArray
from .engine import Engine, bind


@dataclass
class SmoothResult:
    J: np.ndarray            # [n_cfg]
    grad: np.ndarray | None  # [n_cfg][n_params]
    n_capped: int


class SmoothFlux:
    """The objective over a fixed molecule sample on one device.

    ``x, y, vx, vy, eps`` are host arrays (uploaded once); ``sigma_vz`` scales eps into vz.
    ``ckpt_every`` is the lens checkpoint interval K; ``max_steps`` bounds the full RK4 steps per lens
    (default: what the slowest molecule of the evaluated parameter sets needs).

    The molecules are uploaded in order of eps, hence of vz for every parameter set: each wavefront
    of the kernel takes 64 consecutive molecules through every lens step without refill, and
    molecules of nearly equal vz need nearly equal numbers of steps.  J is a mean over the
    molecules, so the order changes nothing but the rounding of the final sums."""

    def __init__(self, beamline, x, y, vx, vy, eps, *, sigma_vz: float, softness: float = 50e-6,
                 z_src: float = 0.0, ckpt_every: int = 8, max_steps: int | None = None, ctx=None, block: int = 0):
        self.engine = Engine(beamline, ctx)
        self.ctx = self.engine.ctx
        self.beamline = beamline
        host = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (x, y, vx, vy, eps)]
        self.n = host[0].size
        if any(a.size != self.n for a in host):
            raise ValueError("source arrays differ in length")
        order = np.argsort(host[4], kind="stable")
        host = [np.ascontiguousarray(a[order]) for a in host]
        self.eps_min = float(host[4].min()) if self.n else 0.0
        self.eps_max = float(host[4].max()) if self.n else 0.0
        self._eps = host[4]
        self.src = [DeviceArray(self.ctx, self.n).upload(a, sync=False) for a in host]
        self.ctx.sync()
        self.sigma_vz = float(sigma_vz)
        self.softness = float(softness)
        self.z_src = float(z_src)
        self.ckpt_every = int(ckpt_every)
        self.max_steps = max_steps
        self.block = int(block)
        npar = C.c_int32()
        theta = np.zeros(N.MAX_ELEMS * 2 + 1)
        N.check("ctraj_smooth_params",
                N.lib().ctraj_smooth_params(self.ctx.handle, C.byref(npar), theta.ctypes.data), self.ctx.handle)
        self.n_params = npar.value
        self._theta_beamline = theta[: self.n_params - 1].copy()
        self._bufs: dict = {}
        self._capped = DeviceArray(self.ctx, 1, np.int64)

    # ------------------------------------------------------------------ parameters
    @property
    def names(self) -> list:
        out = []
        for el in self.beamline.elements:
            if hasattr(el, "V"):
                out += [f"{el.name}.V", f"{el.name}.z0"]
            else:
                out.append(f"{el.name}.r")
        return out + ["mu_vz"]

    def theta0(self, mu_vz: float) -> np.ndarray:
        """The beamline's own parameters (radii, lens V and z0) followed by mu_vz."""
        return np.concatenate([self._theta_beamline, [float(mu_vz)]])

    def steps_needed(self, params: np.ndarray) -> int:
        """Full RK4 steps the slowest molecule needs in the longest lens under these parameter sets."""
        lens = [el for el in self.beamline.elements if hasattr(el, "V")]
        if not lens or self.n == 0:
            return 1
        mus = np.atleast_2d(params)[:, -1]
        lo = self.eps_min if self.sigma_vz >= 0 else self.eps_max
        vz = mus + self.sigma_vz * lo
        if np.any(vz <= 0):
            vz_all = np.add.outer(mus, self.sigma_vz * self._eps)
            vz = vz_all[vz_all > 0]
            if vz.size == 0:
                return 1
        vmin = float(np.min(vz))
        worst = max(el.L / (vmin * el.dt) for el in lens)
        return int(min(max(np.floor(worst) + 1, 1), 1 << 20))

    # ------------------------------------------------------------------ device form
    def launch(self, params_ptr: int, n_cfg: int, J_ptr: int, grad_ptr: int | None, max_steps: int) -> None:
        """Queue one objective (+ adjoint when grad_ptr) launch on device pointers (no synchronisation)."""
        bind(self.ctx, self.engine.cb)
        s = N.Smooth()
        s.n = self.n
        s.x, s.y, s.vx, s.vy, s.eps = (a.ptr for a in self.src)
        s.z_src = self.z_src
        s.sigma_vz = self.sigma_vz
        s.softness = self.softness
        s.n_cfg = int(n_cfg)
        s.n_params = self.n_params
        s.params = params_ptr
        s.J = J_ptr
        s.grad = grad_ptr
        s.n_capped = self._capped.ptr
        s.ckpt_every = self.ckpt_every
        s.max_steps = int(max_steps)
        s.block = self.block
        N.check("ctraj_smooth_objective", N.lib().ctraj_smooth_objective(self.ctx.handle, C.byref(s)),
                self.ctx.handle)

    def _buf(self, name, n, dtype=np.float64):
        b = self._bufs.get(name)
        if b is None or b.n < n:
            b = DeviceArray(self.ctx, max(n, 1), dtype)
            self._bufs[name] = b
        return b

    def evaluate(self, params, grad: bool = True) -> SmoothResult:
        """J (and dJ/dtheta) for params [n_params] or [n_cfg][n_params] (host arrays)."""
        P = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
        if P.shape[1] != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters ({', '.join(self.names)})")
        n_cfg = P.shape[0]
        ms = self.max_steps or self.steps_needed(P)
        dP = self._buf("params", P.size)
        self.ctx.memcpy(dP.ptr, P.ctypes.data, P.nbytes, N.H2D)
        dJ = self._buf("J", n_cfg)
        dG = self._buf("grad", n_cfg * self.n_params) if grad else None
        self._capped.fill(0)
        self.launch(dP.ptr, n_cfg, dJ.ptr, None if dG is None else dG.ptr, ms)
        J = np.empty(n_cfg)
        self.ctx.memcpy(J.ctypes.data, dJ.ptr, J.nbytes, N.D2H)
        G = None
        if grad:
            G = np.empty((n_cfg, self.n_params))
            self.ctx.memcpy(G.ctypes.data, dG.ptr, G.nbytes, N.D2H)
        cap = np.empty(1, np.int64)
        self.ctx.memcpy(cap.ctypes.data, self._capped.ptr, 8, N.D2H)
        self.ctx.sync()
        return SmoothResult(J, G, int(cap[0]))

    def central_differences(self, theta, rel_step: float = 1e-6) -> np.ndarray:
        """Central finite differences of J with h_j = rel_step |theta_j|, all 2 n_params evaluations
        batched into one objective-only launch."""
        theta = np.asarray(theta, dtype=np.float64)
        tp = theta + rel_step * np.abs(theta)
        tm = theta - rel_step * np.abs(theta)
        P = np.repeat(theta[None, :], 2 * self.n_params, axis=0)
        for j in range(self.n_params):
            P[2 * j, j] = tp[j]
            P[2 * j + 1, j] = tm[j]
        J = self.evaluate(P, grad=False).J
        return (J[0::2] - J[1::2]) / (tp - tm)


_FN = None


def _function():
    global _FN
    if _FN is None:
        import torch

        class SmoothFluxFunction(torch.autograd.Function):
            """forward: one objective + adjoint launch; backward: incoming gradient times dJ/dtheta."""

            @staticmethod
            def forward(ctx, th, problem):
                P = th.detach().to("cpu", torch.float64).numpy()
                r = problem.evaluate(P, grad=True)
                if r.n_capped:
                    raise RuntimeError(f"{r.n_capped} lens crossings exceeded max_steps")
                G = torch.from_numpy(r.grad).to(th.device)
                ctx.save_for_backward(G)
                ctx.vector = th.dim() == 1
                J = torch.from_numpy(r.J).to(th.device)
                return J[0] if ctx.vector else J

            @staticmethod
            def backward(ctx, gJ):
                (G,) = ctx.saved_tensors
                if ctx.vector:
                    return gJ * G[0], None
                return gJ[:, None] * G, None

        _FN = SmoothFluxFunction
    return _FN


def smooth_flux(theta, problem: SmoothFlux):
    """J(theta) as a differentiable torch value: a scalar for theta [n_params], a [n_cfg] vector for
    theta [n_cfg][n_params] (many parameter sets in one launch)."""
    return _function().apply(theta, problem)
