"""GPU: the smoothed detected flux and its adjoint kernel (ctraj_smooth_objective).

Checked against the CPU FP64 adjoint of the oracle (same operations; only exp() and the order of
the final sums differ), against central finite differences of the GPU objective itself (the
differentiable workload: 10^6 molecules, relative step 1e-6, agreement 1e-5), and for internal
consistency: batching, determinism, checkpoint interval, the sharp-sigmoid limit against the
ordinary walk, and the torch.autograd.Function.
"""
import numpy as np
import pytest

from oracle import oracle as O
from trajectories import configs
from trajectories import _native as N
from trajectories.beamline import Beamline
from trajectories.beamline_elements import CircularAperture, RectangularAperture
from trajectories.compile import compile_elements
from trajectories.smooth import SmoothFlux, smooth_flux

pytestmark = pytest.mark.gpu


def problem(n, seed=3, sign=-1.0, **kw):
    bl = configs.config1(stark=configs.stark_table(sign))
    src = configs.config3_source(n, seed)
    kw.setdefault("softness", configs.CONFIG3_SOFTNESS)
    kw.setdefault("ckpt_every", configs.CONFIG3_CKPT)
    return bl, src, SmoothFlux(bl, *src, sigma_vz=configs.VZ_SIGMA, **kw)


def scaled_close(g, ref, theta, rtol, atol_frac):
    gs, rs = np.asarray(g) * theta, np.asarray(ref) * theta
    scale = np.max(np.abs(rs))
    err = np.abs(gs - rs) / np.maximum(np.abs(rs), atol_frac * scale / rtol)
    print("theta*dJ/dtheta:", gs, "\nreference:      ", rs, "\nscaled error:", err)
    assert np.all(err <= rtol), err


@pytest.mark.parametrize("sign", [-1.0, +1.0])
def test_adjoint_matches_cpu_fp64_adjoint(sign):
    bl, src, pr = problem(20_000, sign=sign)
    theta = configs.config3_theta()
    r = pr.evaluate(theta)
    o = O.smooth(pr.engine.cb, *src, theta, sigma_vz=configs.VZ_SIGMA, softness=configs.CONFIG3_SOFTNESS,
                 max_steps=pr.steps_needed(theta))
    print("J gpu", r.J[0], "cpu", o["J"][0])
    assert r.J[0] == pytest.approx(o["J"][0], rel=1e-12, abs=1e-300)
    scaled_close(r.grad[0], o["grad"][0], theta, rtol=1e-10, atol_frac=1e-12)


def test_config3_gradient_matches_central_differences():
    """The differentiable workload: 10^6 molecules; every gradient component that is not negligible
    (above 1e-8 of the largest, in theta_j dJ/dtheta_j) agrees with central differences of step
    1e-6 |theta_j| to 1e-5 relative."""
    bl, src, pr = problem(1_000_000)
    theta = configs.config3_theta()
    r = pr.evaluate(theta)
    fd = pr.central_differences(theta, 1e-6)
    print("names", pr.names, "J", r.J[0])
    scaled_close(r.grad[0], fd, theta, rtol=1e-5, atol_frac=1e-8)


def test_objective_only_and_checkpoint_interval_do_not_change_results():
    """The forward of the adjoint launch is the objective-only launch, and the checkpoint interval
    only changes what is recomputed: J and the gradient are bitwise identical for K = 1, 7, 64, 500."""
    theta = configs.config3_theta()
    res = []
    for K in (1, 7, 64, 500):
        _, _, pr = problem(5_000, ckpt_every=K)
        res.append(pr.evaluate(theta))
        assert pr.evaluate(theta, grad=False).J[0] == res[-1].J[0]
    for r in res[1:]:
        assert r.J[0] == res[0].J[0]
        np.testing.assert_array_equal(r.grad, res[0].grad)


def test_batched_parameter_sets_and_determinism():
    """A launch over several parameter sets equals the sets launched one by one, bit for bit, and a
    repeated launch reproduces itself (fixed-order reductions)."""
    _, _, pr = problem(30_000)
    t = configs.config3_theta()
    P = np.stack([t, t * [1, 0.7, 1, 1, 1, 1], t * [1.2, 1, 1.05, 0.9, 1.1, 1.0], t * [1, 1, 1, 1, 1, 0.95]])
    rb = pr.evaluate(P)
    for c in range(P.shape[0]):
        rs = pr.evaluate(P[c])
        assert rs.J[0] == rb.J[c]
        np.testing.assert_array_equal(rs.grad[0], rb.grad[c])
    rb2 = pr.evaluate(P)
    np.testing.assert_array_equal(rb.J, rb2.J)
    np.testing.assert_array_equal(rb.grad, rb2.grad)


def test_sharp_limit_counts_the_detected_molecules_of_the_walk():
    """softness -> 0 turns J into the fraction of molecules inside every aperture: n J equals the
    detected count of the ordinary walker on the same molecules with the lens bore opened, except
    for molecules within 1e-9 m of an aperture edge."""
    n = 100_000
    bl, src, pr = problem(n, seed=9, sign=+1.0, softness=1e-15)
    theta = configs.config3_theta()
    J = pr.evaluate(theta, grad=False).J[0]
    cb = compile_elements(bl.elements, bl.g)
    cb.elems[1].p[0] = 1e9
    cb.elems[1].p[1] = 1e18
    x, y, vx, vy, eps = src
    vz = configs.VZ_MEAN + configs.VZ_SIGMA * eps
    o = O.propagate(cb, x, y, np.zeros(n), vx, vy, vz)
    det = int(np.count_nonzero(o["label"] == cb.label_detected))
    n_edge = int(np.count_nonzero(o["edge"] < 1e-9))
    print("n J", J * n, "detected", det, "edge molecules", n_edge)
    assert det > 100
    assert abs(J * n - det) <= n_edge + 1e-6


def test_autograd_function_returns_the_adjoint_gradient():
    import torch

    _, _, pr = problem(10_000)
    t = configs.config3_theta()
    ref = pr.evaluate(t)
    th = torch.tensor(t, dtype=torch.float64, requires_grad=True)
    J = smooth_flux(th, pr)
    assert J.item() == ref.J[0]
    (3.0 * J).backward()
    np.testing.assert_array_equal(th.grad.numpy(), 3.0 * ref.grad[0])
    # a batch of parameter sets: one launch, one vector of J
    P = torch.tensor(np.stack([t, t * [1, 1.1, 1, 1, 1, 1]]), dtype=torch.float64, requires_grad=True)
    Jv = smooth_flux(P, pr)
    Jv.sum().backward()
    np.testing.assert_array_equal(P.grad[0].numpy(), ref.grad[0])
    if torch.cuda.is_available():
        thc = torch.tensor(t, dtype=torch.float64, device="cuda", requires_grad=True)
        Jc = smooth_flux(thc, pr)
        Jc.backward()
        np.testing.assert_array_equal(thc.grad.cpu().numpy(), ref.grad[0])


def test_capped_lens_crossings_are_reported_and_bad_arguments_rejected():
    _, _, pr = problem(2_000, max_steps=10)
    t = configs.config3_theta()
    assert pr.evaluate(t).n_capped > 0
    import torch

    with pytest.raises(RuntimeError):
        smooth_flux(torch.tensor(t, dtype=torch.float64, requires_grad=True), pr)
    _, _, pr2 = problem(2_000)
    with pytest.raises(ValueError):
        pr2.evaluate(t[:-1])
    bl = Beamline([RectangularAperture("box", 0.3, 0.0, w=1e-2, h=1e-2), CircularAperture("det", 1.0, 0.0, d=1e-2)])
    with pytest.raises(N.NativeError):
        SmoothFlux(bl, *configs.config3_source(100), sigma_vz=16.0)
