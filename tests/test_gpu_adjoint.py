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


def test_gradient_fodo48_full_size():
    """The adjoint workload at full size: 48-cell FODO, 65 536 particles
    (sigma_x = sigma_y = 3 mm), 512 turns, checkpoints every 32, amplitude
    loss with the soft-survival penalty; the gradient with respect to 16 of
    the 192 strengths (spread over the four families) against central
    differences with h = 1e-7, within 1e-6 of the largest component."""
    cells, turns = 48, 512
    ring, shared, theta, bunch = _fodo_setup(cells, 65536)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    tr = ufo.AdjointTracker(ring, flags, turns, COORDS, shared, every=32)
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
    idx = [f * cells + c for f in range(4) for c in (0, 17, 31, 47)]
    eye = np.eye(len(theta))
    fd = np.array([(L(theta + h * eye[i]) - L(theta - h * eye[i])) / (2 * h) for i in idx])
    err = np.abs(g[idx] - fd) / np.max(np.abs(fd))
    print("max rel err vs central differences:", err.max())
    assert err.max() < 1e-6


@pytest.mark.parametrize("cells,turns,n,every", [(8, 48, 256, 16), (48, 512, 65536, 32)])
def test_initial_coordinate_adjoint_matches_tangent_jacobian(cells, turns, n, every):
    """dL/dx0 from the reverse pass equals J^T dL/dx_T with J from the
    forward-tangent kernel of the same lattice (parameters baked in); also
    on the full adjoint workload (FODO-48, 65 536 particles, 512 turns)."""
    ring, shared, theta, bunch = _fodo_setup(cells, n, seed=1)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    tr = ufo.AdjointTracker(ring, flags, turns, COORDS, shared, every=every)
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


def _every_kind_ring(cells=4):
    """A stable ring (KICK tunes 0.62 / 0.73) whose cell holds every element
    kind: teapot quadrupoles with multipole errors and a misalignment, a
    combined-function bend with pole faces and fringe, sextupole, octupole,
    thin multipole to order 4 with skew terms, wire; aperture at 20 mm."""
    qf = ufo.Quadrupole("QF", length=0.5, k1=1.0, slices=4, dx=2e-5, dy=-1e-5, dkn=[0, 0, 3.0], dks=[0, 0.01])
    qd = ufo.Quadrupole("QD", length=0.5, k1=-1.0, slices=4)
    b = ufo.Sbend("B", length=0.8, angle=0.05, k1=-0.05, e1=0.02, e2=0.03, hgap=0.02, fint=0.5, slices=4)
    s = ufo.Sextupole("S", length=0.2, k2=5.0, dx=1e-5)
    o = ufo.Octupole("O", length=0.2, k3=50.0, slices=2)
    m = ufo.Multipole("M", knl=[0.0, 0.001, 0.5, 10.0, 100.0], ksl=[0.0, 0.001, 0.2, 5.0, 0.0])
    w = ufo.Wire("W", x=0.05, y=0.03, k=1e-7)
    d = ufo.Drift("D", length=0.5)
    dt = ufo.Drift("DT", length=0.3)
    cell = ufo.Line("C", [qf, d, b, d, s, qd, d, o, m, d, w, dt])
    ap = ufo.Aperture("AP", window="(fabs(x) < 0.02 && fabs(y) < 0.02)")
    return ufo.Line("RING", [cell] * cells + [ap])


@pytest.mark.parametrize("exact", [False, True])
def test_gradient_every_element_kind(exact):
    """Gradient through a ring of every element kind with respect to two
    quadrupole strengths, sextupole and octupole strengths, a multipole
    coefficient and a misalignment, against central differences; with EXACT
    the drifts are the square-root form.  About 3 % of the bunch is lost."""
    line = _every_kind_ring()
    shared = [("QF", "k1"), ("QD", "k1"), ("S", "k2"), ("O", "k3"), ("M", "k3"), ("QF", "dx")]
    theta = np.array([1.0, -1.0, 5.0, 50.0, 10.0, 2e-5])
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | (ufo.EXACT if exact else 0)
    turns = 30
    bunch = lattices.gaussian_bunch(1024, seed=4, sx=4e-3, sy=4e-3)
    tr = ufo.AdjointTracker(line, flags, turns, COORDS, shared, every=8)
    dev = tr.device
    th = torch.tensor(theta, device=dev, requires_grad=True)
    x0 = torch.tensor(bunch, device=dev)
    xf, lost = adjoint.track(tr, th, x0)
    loss = adjoint.amplitude_loss(xf, aperture=0.02)
    loss.backward()
    g = th.grad.cpu().numpy()

    def L(t):
        with torch.no_grad():
            xf_, _ = tr.forward(x0, torch.tensor(t, device=dev))
            return adjoint.amplitude_loss(xf_, aperture=0.02).item()

    fd = []
    for i in range(len(theta)):
        h = 1e-6 * max(abs(theta[i]), 1e-3)
        e = np.zeros_like(theta)
        e[i] = h
        fd.append((L(theta + e) - L(theta - e)) / (2 * h))
    fd = np.array(fd)
    # relative to each parameter's own derivative, floored at 1e-6 of the
    # largest sensitivity |dL/dtheta_j * theta_j| brought to this parameter's scale
    scale = np.max(np.abs(fd * theta)) / np.abs(theta)
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-6 * scale)
    print(f"lost {int((lost < turns).sum())} of {len(bunch)}")
    for i, p in enumerate(shared):
        print(f"{p}: adjoint {g[i]:+.10e} fd {fd[i]:+.10e} rel {rel[i]:.2e}")
    assert 0 < int((lost < turns).sum()) < len(bunch) // 10
    assert np.all(rel < 1e-5)


def test_argument_shapes_are_checked_before_launch():
    """Inputs whose shapes do not match what the kernels index raise on the
    host: x0 columns, the number of strengths, lam, backward before
    forward (single-lattice and batched trackers)."""
    from ufo_amd.batch_adjoint import BatchedAdjoint
    ring, shared, theta, bunch = _fodo_setup(2, 64)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    tr = ufo.AdjointTracker(ring, flags, 8, COORDS, shared, every=4)
    dev = tr.device
    with pytest.raises(RuntimeError):
        tr.backward(torch.zeros((64, 4), dtype=torch.float64, device=dev))
    with pytest.raises(ValueError):
        tr.forward(torch.tensor(bunch[:, :3], device=dev), theta)
    with pytest.raises(ValueError):
        tr.forward(torch.tensor(bunch, device=dev), theta[:-1])
    xf, lost = tr.forward(torch.tensor(bunch, device=dev), theta)
    with pytest.raises(ValueError):
        tr.backward(torch.zeros((63, 4), dtype=torch.float64, device=dev))
    gx0, gth = tr.backward(torch.ones((64, 4), dtype=torch.float64, device=dev))
    assert gth.shape == (len(theta),) and torch.isfinite(gth).all()

    ring, prm, vals, counts, fam = lattices.dba_family_model(seeds=2, count_range=(10, 20))
    with pytest.raises(ValueError):
        BatchedAdjoint(ring, flags, 8, COORDS, prm, vals[:, :-1], counts, fam[:2])
    ba = BatchedAdjoint(ring, flags, 8, COORDS, prm, vals, counts, fam[:2])
    with pytest.raises(RuntimeError):
        ba.backward(torch.zeros((ba.particles, 4), dtype=torch.float64, device=dev))
    with pytest.raises(ValueError):
        ba.forward(torch.zeros((ba.particles + 1, 5), dtype=torch.float64, device=dev))


def _thick_ring(cells=4):
    """A ring of thick elements (tunes 0.431 / 0.358 per the oracle's
    Optics): thick quadrupoles and combined-function sector bends with pole
    faces and fringe."""
    out = []
    for i in range(cells):
        out += [ufo.Quadrupole(f"QF{i}", length=0.4, k1=1.2), ufo.Drift(f"DA{i}", length=1.5),
                ufo.Sbend(f"B{i}", length=1.0, angle=0.1, k1=-0.2, e1=0.05, e2=0.05, hgap=0.02, fint=0.5),
                ufo.Drift(f"DB{i}", length=1.0), ufo.Quadrupole(f"QD{i}", length=0.4, k1=-1.4),
                ufo.Drift(f"DC{i}", length=1.5)]
    return ufo.Line("R", out)


def test_gradient_thick_elements():
    """Thick (matrix) quadrupoles and bends in 6D: the gradient with respect
    to every quadrupole k1, every bend k1 and one bend's angle (through its
    matrices, dispersion and pole faces) against central differences."""
    cells, turns = 4, 40
    ring = _thick_ring(cells)
    shared = ([(f"QF{i}", "k1") for i in range(cells)] + [(f"QD{i}", "k1") for i in range(cells)]
              + [(f"B{i}", "k1") for i in range(cells)] + [("B0", "angle")])
    theta = np.array([1.2] * cells + [-1.4] * cells + [-0.2] * cells + [0.1])
    bunch = lattices.gaussian_bunch(2048, seed=3, sx=2e-3, sy=2e-3)
    flags = ufo.DOUBLE_PRECISION
    tr = ufo.AdjointTracker(ring, flags, turns, COORDS, shared, every=8)
    dev = tr.device
    th = torch.tensor(theta, device=dev, requires_grad=True)
    x0 = torch.tensor(bunch, device=dev)
    xf, lost = adjoint.track(tr, th, x0)
    loss = adjoint.amplitude_loss(xf)
    loss.backward()
    g = th.grad.cpu().numpy()
    assert int((lost < turns).sum()) == 0

    def L(t):
        with torch.no_grad():
            xf_, _ = tr.forward(x0, torch.tensor(t, device=dev))
            return adjoint.amplitude_loss(xf_).item()

    h = 1e-7
    eye = np.eye(len(theta))
    fd = np.array([(L(theta + h * eye[i]) - L(theta - h * eye[i])) / (2 * h) for i in range(len(theta))])
    err = np.abs(g - fd) / np.max(np.abs(fd))
    print("max rel err vs central differences:", err.max(), "|g|", np.abs(g).max())
    assert err.max() < 1e-6


def test_backward_uses_the_buffers_of_its_own_forward():
    """Two differentiable passes on one tracker (bunch A at theta, then
    bunch B of the same size at another theta) before either backward: each
    gradient equals that of its pass differentiated alone (the reverse pass
    reads the checkpoints and coefficients of the forward that built its
    graph, not the tracker's latest).  The same for BatchedAdjoint with
    explicit saved buffers."""
    cells, turns = 4, 32
    ring, shared, theta, bunch_a = _fodo_setup(cells, 1024, seed=1)
    bunch_b = lattices.gaussian_bunch(1024, seed=2, sx=3e-3, sy=3e-3)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    tr = ufo.AdjointTracker(ring, flags, turns, COORDS, shared, every=8)
    dev = tr.device

    def grad(theta_, bunch_):
        th = torch.tensor(theta_, device=dev, requires_grad=True)
        x0 = torch.tensor(bunch_, device=dev, requires_grad=True)
        xf, _ = adjoint.track(tr, th, x0)
        return adjoint.amplitude_loss(xf), th, x0

    la, tha, xa = grad(theta, bunch_a)
    la.backward()
    alone_a = (tha.grad.clone(), xa.grad.clone())
    lb, thb, xb = grad(theta * 1.01, bunch_b)
    lb.backward()
    alone_b = (thb.grad.clone(), xb.grad.clone())
    la, tha, xa = grad(theta, bunch_a)
    lb, thb, xb = grad(theta * 1.01, bunch_b)
    la.backward()                                 # the older graph first
    lb.backward()
    for (g_th, g_x), (a_th, a_x) in (((tha.grad, xa.grad), alone_a), ((thb.grad, xb.grad), alone_b)):
        assert torch.allclose(g_th, a_th, rtol=1e-12, atol=0) and torch.allclose(g_x, a_x, rtol=1e-12, atol=0)
    assert not torch.allclose(alone_a[0], alone_b[0], rtol=1e-3)

    from ufo_amd.batch_adjoint import BatchedAdjoint
    bring, prm, vals, counts, fam = lattices.dba_family_model(seeds=2, count_range=(100, 200))
    ba = BatchedAdjoint(bring, flags, 16, COORDS, prm, vals, counts, fam[:2])
    n = int(counts.sum())
    rng = np.random.default_rng(3)
    x0s = [np.stack([rng.normal(0, s, n), np.zeros(n), rng.normal(0, s / 2, n), np.zeros(n), np.zeros(n)], 1)
           for s in (2e-3, 4e-3)]
    lam = torch.ones((n, 4), dtype=torch.float64, device=dev)
    alone = []
    for x0 in x0s:
        ba.forward(x0)
        alone.append(ba.backward(lam)[1].clone())
    _, _, saved_a = ba.forward_saved(x0s[0])
    ba.forward(x0s[1])
    ga = ba.backward(lam, saved_a)[1]
    assert torch.allclose(ga, alone[0], rtol=1e-12, atol=0)
    assert not torch.allclose(alone[0], alone[1], rtol=1e-3)
