"""The adjoint kernels (``adjoint.adjoint_kernels``) built for the host
(one thread per lane, workgroup barriers and FP64 atomics emulated,
``host_emulation.HostAdjoint``): the reverse pass's dL/dtheta and dL/dx0
against central differences of the forward pass, for every element kind the
adjoint differentiates, on the CPU.  The GPU tests check the same kernels on
the device at full size."""
import numpy as np
import pytest

import ufo_amd as ufo
from ufo_amd import lattices

from host_emulation import HostAdjoint

COORDS = ["x", "px", "y", "py", "dp"]


def _check(line, flags, shared, theta, bunch, turns=24, tol=1e-6):
    adj = HostAdjoint(line, flags, turns, COORDS, shared, every=8)
    rng = np.random.default_rng(7)
    w = rng.normal(size=(len(bunch), 4))
    xf, lost, saved = adj.forward(bunch, theta)
    assert np.all(lost == turns)

    def loss(t, x0=bunch):
        return float(np.sum(w * adj.forward(x0, t)[0]))

    gx0, g = adj.backward(w, saved)
    # central differences with a step relative to each parameter's size
    h = 1e-6 * np.maximum(np.abs(theta), 1e-2)
    eye = np.eye(len(theta))
    fd = np.array([(loss(theta + h[i] * eye[i]) - loss(theta - h[i] * eye[i])) / (2 * h[i])
                   for i in range(len(theta))])
    err = np.abs(g - fd) / np.max(np.abs(fd))
    # dL/dx0 of three particles, every coordinate
    gx_fd = np.zeros((3, 4))
    for p in range(3):
        for c in range(4):
            hx = 1e-8
            up, dn = bunch.copy(), bunch.copy()
            up[p, c] += hx
            dn[p, c] -= hx
            gx_fd[p, c] = (loss(theta, up) - loss(theta, dn)) / (2 * hx)
    err_x = np.abs(gx0[:3] - gx_fd) / np.max(np.abs(gx_fd))
    print(f"dL/dtheta vs central differences: max rel {err.max():.2e}; dL/dx0: {err_x.max():.2e}")
    assert err.max() < tol
    assert err_x.max() < tol
    return g


def test_fodo_kick_strengths_and_misalignment():
    """Teapot quadrupoles (k1), thin sextupole multipoles (knl[2] as k2, the
    14-operation normal-sextupole reverse) and a quadrupole misalignment."""
    ring = lattices.fodo_ring(4, distinct=True)
    shared = [("QF0", "k1"), ("QD1", "k1"), ("SF2", "k2"), ("SD3", "k2"), ("QF1", "dx"), ("QD2", "dy")]
    theta = np.array([1.2, -1.2, 0.8, -1.4, 2e-4, -1e-4])
    bunch = lattices.gaussian_bunch(40, seed=1, sx=3e-3, sy=3e-3)
    _check(ring, ufo.KICK | ufo.DOUBLE_PRECISION, shared, theta, bunch)


def test_thick_elements():
    """Thick quadrupoles and combined-function bends (matrices, dispersion,
    pole faces and fringe): quadrupole and bend k1 and a bend angle."""
    out = []
    for i in range(2):
        out += [ufo.Quadrupole(f"QF{i}", length=0.4, k1=1.2), ufo.Drift(f"DA{i}", length=1.5),
                ufo.Sbend(f"B{i}", length=1.0, angle=0.1, k1=-0.2, e1=0.05, e2=0.05, hgap=0.02, fint=0.5),
                ufo.Drift(f"DB{i}", length=1.0), ufo.Quadrupole(f"QD{i}", length=0.4, k1=-1.4),
                ufo.Drift(f"DC{i}", length=1.5)]
    ring = ufo.Line("R", out)
    shared = [("QF0", "k1"), ("QD1", "k1"), ("B0", "k1"), ("B1", "angle")]
    theta = np.array([1.2, -1.4, -0.2, 0.1])
    bunch = lattices.gaussian_bunch(40, seed=2, sx=2e-3, sy=2e-3)
    _check(ring, ufo.DOUBLE_PRECISION, shared, theta, bunch)


def test_skew_and_higher_order_kicks():
    """A multipole with normal and skew entries up to order 3 (the Horner
    reverse), a teapot sextupole with a skew error (the two-part quadratic
    reverse) and an octupole, each strength a parameter."""
    cell = [ufo.Quadrupole("QF", length=0.5, k1=1.2), ufo.Drift("D1", length=1.0),
            ufo.Multipole("M", knl=[0.0, 0.0, 0.6, 5.0], ksl=[0.0, 0.05, 0.3]),
            ufo.Sextupole("S", length=0.2, k2=4.0, dks=[0.0, 0.0, 0.5]), ufo.Drift("D2", length=1.0),
            ufo.Quadrupole("QD", length=0.5, k1=-1.2), ufo.Octupole("O", length=0.2, k3=30.0),
            ufo.Drift("D3", length=2.0)]
    ring = ufo.Line("R", cell * 3)
    shared = [("M", "k2"), ("M", "k3"), ("M", "k1s"), ("M", "k2s"), ("S", "k2"), ("S", "dk2s"), ("O", "k3")]
    theta = np.array([0.6, 5.0, 0.05, 0.3, 4.0, 0.5, 30.0])
    bunch = lattices.gaussian_bunch(40, seed=3, sx=2e-3, sy=2e-3)
    _check(ring, ufo.KICK | ufo.DOUBLE_PRECISION, shared, theta, bunch)


def test_batched_family_gradients_per_seed():
    """The batched adjoint (``BatchedAdjoint``'s kernels) on the host: two
    seeds of the error model with uneven particle counts (several tiles,
    partly filled), the per-seed gradient with respect to all 8 sextupole
    families (both period sets, so the set-B rows of ``ufo_gmap`` route
    gradients) against central differences of each seed's own loss."""
    from ufo_amd.batch_adjoint import BatchedAdjoint
    from host_emulation import HostBatchedAdjoint
    ring, prm, vals, counts, fam = lattices.dba_family_model(seeds=2, count_range=(70, 140))
    turns = 16
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    rng = np.random.default_rng(5)
    n = int(counts.sum())
    x0 = np.stack([rng.normal(0, 3e-3, n), rng.normal(0, 1e-4, n), rng.normal(0, 1.5e-3, n),
                   rng.normal(0, 1e-4, n), rng.normal(0, 1e-3, n)], 1)
    seed_of = np.repeat(np.arange(len(counts)), counts)
    w = rng.normal(size=(n, 4))

    def run(values):
        return HostBatchedAdjoint(BatchedAdjoint(ring, flags, turns, COORDS, prm, values, counts, fam))

    host = run(vals)
    assert host.b.kbwd.src.count("ufo_gmap[rep]") > 0
    xf, lost, saved = host.forward(x0)
    assert np.all(lost == turns)
    _, gseed = host.backward(w, saved)

    def per_seed_loss(state):
        return np.array([np.sum((w * state)[seed_of == s]) for s in range(len(counts))])

    col = [prm.index(f) for f in fam]
    fd = np.zeros_like(gseed)
    for j, c in enumerate(col):
        h = 1e-6 * max(np.abs(vals[:, c]).max(), 1e-2)
        vp, vm = vals.copy(), vals.copy()
        vp[:, c] += h
        vm[:, c] -= h
        fd[:, j] = (per_seed_loss(run(vp).forward(x0)[0]) - per_seed_loss(run(vm).forward(x0)[0])) / (2 * h)
    err = np.abs(gseed - fd) / np.abs(fd).max()
    print(f"per-seed family gradients vs central differences: max rel {err.max():.2e}")
    assert err.max() < 1e-6


@pytest.mark.parametrize("mode", ["6d", "fived"])
def test_every_cell_strength_rotated_loop(mode):
    """A ring of 6 FODO cells with every quadrupole k1 and sextupole k2 a
    parameter of its own (the per-cell strengths of the 48-cell workload):
    the reverse pass loops over the cells with the loop body rotated past
    the leading drift (the drift closing one cell and the one opening the
    next are one drift), runs on scaled momenta p (1 + dp) so the kicks use
    the table's strengths without a product with 1/(1+dp), and leaves the
    constant slice factors to the reduction (each parameter's total is
    scaled once)."""
    from ufo_amd import adjoint
    cells = 6
    ring = lattices.fodo_ring(cells, distinct=True)
    shared = ([(f"QF{i}", "k1") for i in range(cells)] + [(f"QD{i}", "k1") for i in range(cells)]
              + [(f"SF{i}", "k2") for i in range(cells)] + [(f"SD{i}", "k2") for i in range(cells)])
    theta = np.array([1.2] * cells + [-1.2] * cells + [0.8] * cells + [-1.4] * cells)
    theta = theta * (1.0 + 0.01 * np.arange(len(theta)) / len(theta))
    bunch = lattices.gaussian_bunch(40, seed=4, sx=3e-3, sy=3e-3)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    coords = COORDS
    if mode == "fived":
        flags |= ufo.FIVED
        coords = ["x", "px", "y", "py"]
        bunch = bunch[:, :4]
    fk, bk, tk, layout, _ = adjoint.adjoint_kernels(ring, flags, 24, coords, shared, dp=1e-3, every=8)
    assert "for (int rep = 4; rep >= 0; --rep)" in bk.src        # 6 cells: 5 in the loop
    assert sorted(set(layout["gscale"])) != [1.0]
    loop = bk.src.split("for (int rep = 4; rep >= 0; --rep)")[1]
    assert "px = c[np] / oodppo" in bk.src and "* oodppo" not in loop
    if mode == "fived":
        adj = HostAdjoint(ring, flags, 24, coords, shared, every=8, dp=1e-3)
        assert adj.layout["gscale"] == layout["gscale"]
        rng = np.random.default_rng(7)
        w = rng.normal(size=(len(bunch), 4))
        xf, lost, saved = adj.forward(bunch, theta)
        assert np.all(lost == 24)
        gx0, g = adj.backward(w, saved)
        h = 1e-6 * np.abs(theta)
        eye = np.eye(len(theta))
        fd = np.array([(np.sum(w * adj.forward(bunch, theta + h[i] * eye[i])[0])
                        - np.sum(w * adj.forward(bunch, theta - h[i] * eye[i])[0])) / (2 * h[i])
                       for i in range(len(theta))])
        err = np.abs(g - fd) / np.max(np.abs(fd))
        print(f"FIVED dL/dtheta vs central differences: max rel {err.max():.2e}")
        assert err.max() < 1e-6
    else:
        _check(ring, flags, shared, theta, bunch)


def test_family_strength_over_magnets_of_different_lengths():
    """One sextupole family strength bound (by label) to magnets of two
    lengths in every cell, plus per-cell quadrupole strengths: the family's
    sums carry different constant factors (k2 L / (2 n) per slice), so the
    reverse pass multiplies them at each flush while the quadrupoles' totals
    are scaled after the reduction."""
    from ufo_amd import adjoint
    cells = 4
    out = []
    for i in range(cells):
        out += [ufo.Quadrupole(f"QF{i}", length=0.5, k1=1.2), ufo.Sextupole("SF", length=0.2, k2=6.0),
                ufo.Drift(f"DA{i}", length=2.0), ufo.Quadrupole(f"QD{i}", length=0.5, k1=-1.2),
                ufo.Sextupole("SF", length=0.3, k2=6.0), ufo.Drift(f"DB{i}", length=2.0)]
    ring = ufo.Line("R", out)
    shared = [(f"QF{i}", "k1") for i in range(cells)] + [("SF", "k2")]
    theta = np.array([1.2, 1.21, 1.19, 1.2, 6.0])
    fk, bk, tk, layout, _ = adjoint.adjoint_kernels(ring, ufo.KICK | ufo.DOUBLE_PRECISION, 24, COORDS, shared,
                                                    every=8)
    assert layout["gscale"][-1] == 1.0 and set(layout["gscale"][:cells]) == {0.5 / 8}
    bunch = lattices.gaussian_bunch(40, seed=5, sx=2e-3, sy=2e-3)
    _check(ring, ufo.KICK | ufo.DOUBLE_PRECISION, shared, theta, bunch)
