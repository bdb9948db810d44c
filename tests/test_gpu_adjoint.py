"""GPU checks of the reverse-mode adjoint: gradients against central finite
differences of the forward tracker, initial-coordinate adjoints against the
forward-tangent Jacobian, and the forward state against Track."""
import numpy as np
import pytest
import torch

import ufo_amd as ufo
from ufo_amd import adjoint, lattices

pytestmark = pytest.mark.gpu

COORDS = ["x", "px", "y", "py", "dp"]


def _fodo_setup(cells, n, seed=0, sigma=3e-3):
    ring = lattices.fodo_ring(cells, distinct=True)
    shared = ([(f"QF{i}", "k1") for i in range(cells)] + [(f"QD{i}", "k1") for i in range(cells)]
              + [(f"SF{i}", "k2") for i in range(cells)] + [(f"SD{i}", "k2") for i in range(cells)])
    theta = np.array([1.2] * cells + [-1.2] * cells + [0.8] * cells + [-1.4] * cells)
    rng = np.random.default_rng(seed)
    theta = theta * (1.0 + 0.01 * rng.normal(size=theta.shape))
    bunch = lattices.gaussian_bunch(n, seed=seed, sx=sigma, sy=sigma)
    return ring, shared, theta, bunch


def _bake(ring, shared, theta):
    """The same ring with theta written into the elements."""
    flat = {e.label: e for e in ring.flatten()}
    for (label, name), v in zip(shared, theta):
        setattr(flat[label], name, float(v))
    return ring


def test_gradient_against_central_differences():
    cells, turns = 8, 64
    ring, shared, theta, bunch = _fodo_setup(cells, 2048)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    tr = ufo.AdjointTracker(ring, flags, turns, COORDS, shared, every=16)
    dev = tr.device
    th = torch.tensor(theta, device=dev, requires_grad=True)
    x0 = torch.tensor(bunch, device=dev)
    xf, lost = adjoint.track(tr, th, x0)
    loss = adjoint.amplitude_loss(xf)
    loss.backward()
    g = th.grad.cpu().numpy()
    print(f"loss {loss.item():.6e}, lost {int((lost < turns).sum())}, |g| max {np.abs(g).max():.3e}")

    def L(t):
        with torch.no_grad():
            xf_, _ = tr.forward(x0, torch.tensor(t, device=dev))
            return adjoint.amplitude_loss(xf_).item()

    h = 1e-7
    idx = list(range(0, len(theta), 5))
    fd = np.array([(L(theta + h * np.eye(len(theta))[i]) - L(theta - h * np.eye(len(theta))[i])) / (2 * h)
                   for i in idx])
    err = np.abs(g[idx] - fd) / np.max(np.abs(fd))
    print("max rel err vs central differences:", err.max())
    assert err.max() < 1e-6


def test_initial_coordinate_adjoint_matches_tangent_jacobian():
    """dL/dx0 from the reverse pass equals J^T dL/dx_T with J from the
    forward-tangent kernel of the same lattice (parameters baked in)."""
    cells, turns = 8, 48
    ring, shared, theta, bunch = _fodo_setup(cells, 256, seed=1)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    tr = ufo.AdjointTracker(ring, flags, turns, COORDS, shared, every=16)
    dev = tr.device
    th = torch.tensor(theta, device=dev)
    x0 = torch.tensor(bunch, device=dev, requires_grad=True)
    xf, lost = adjoint.track(tr, th, x0)
    xf.retain_grad()
    loss = adjoint.amplitude_loss(xf)
    loss.backward()
    lam = xf.grad.cpu().numpy()
    gx0 = x0.grad.cpu().numpy()[:, :4]
    baked = _bake(ring, shared, theta)
    tm = ufo.TangentMap(baked, flags=flags, turns=turns, particles=len(bunch), parameters=COORDS)
    tm.parameters[:] = bunch
    tm.run()
    alive = (lost.cpu().numpy() == turns)
    # the adjoint's forward is the tracker's arithmetic: same final state
    assert np.array_equal(xf.detach().cpu().numpy()[alive], tm.state[alive])
    expect = np.einsum("pij,pi->pj", tm.jacobian, lam)
    scale = np.max(np.abs(expect[alive]))
    err = np.max(np.abs(gx0[alive] - expect[alive])) / scale
    print("dL/dx0 vs J^T lambda:", err)
    assert err < 1e-9
    assert np.all(gx0[~alive] == 0)
