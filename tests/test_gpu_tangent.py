"""GPU checks of the forward-tangent kernel and the Newton closed orbit."""
import json
import os

import numpy as np
import pytest

import ufo_amd as ufo
from ufo_amd import lattices
from ufo_amd.co import symplectic_error
from oracle import runner

from conftest import OPTICS_DIR, ROOT

pytestmark = pytest.mark.gpu
G = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_values.json")))
C4 = ["x", "px", "y", "py"]


def _fd_jacobian(line, x0, flags, turns, h, dp=0.0):
    """Central differences of the oracle's N-turn map [particles, 4, 4]."""
    n = len(x0)
    J = np.zeros((n, 4, 4))
    for j in range(4):
        step = np.zeros(4)
        step[j] = h[j]
        a = runner.track(line, x0 + step, flags=flags, turns=turns, parameters=C4, where=[-1], dp=dp)[:, -1, :4]
        b = runner.track(line, x0 - step, flags=flags, turns=turns, parameters=C4, where=[-1], dp=dp)[:, -1, :4]
        J[:, :, j] = (a - b) / (2 * h[j])
    return J


@pytest.mark.parametrize("case", ["fodo_kick", "alba_thick", "alba_kick"])
def test_tangent_map_against_finite_differences(case):
    if case == "fodo_kick":
        line, flags, turns, dp = lattices.fodo_ring(16), ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED, 3, 0.0
    else:
        line = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad")).RING
        flags = ufo.DOUBLE_PRECISION | ufo.FIVED | (ufo.KICK if case == "alba_kick" else 0)
        turns, dp = 1, 1e-3
    rng = np.random.default_rng(4)
    n = 16
    x0 = np.stack([rng.normal(0, 2e-3, n), rng.normal(0, 1e-4, n), rng.normal(0, 1e-3, n),
                   rng.normal(0, 1e-4, n)], 1)
    tm = ufo.TangentMap(line, flags=flags, turns=turns, particles=n, parameters=C4, dp=dp)
    tm.parameters[:] = x0
    tm.run()
    # the state path of the tangent kernel is the forward kernel's arithmetic
    tr = ufo.Track(line, flags=flags, turns=turns, particles=n, parameters=C4, where=[-1], dp=dp)
    tr.parameters[:] = x0
    tr.run()
    assert np.array_equal(tm.state, tr.tracks[:, -1, :4])
    J = tm.jacobian
    Jfd = _fd_jacobian(line, x0, flags, turns, h=[1e-7, 1e-8, 1e-7, 1e-8], dp=dp)
    scale = np.max(np.abs(Jfd), axis=(1, 2), keepdims=True)
    err = np.max(np.abs(J - Jfd) / scale)
    print(f"{case}: max |J - J_fd| / max|J| = {err:.2e}; symplectic error {symplectic_error(J).max():.2e}")
    assert err < 1e-6
    assert np.all(symplectic_error(J) < 1e-12)
    assert np.all(np.abs(np.linalg.det(J) - 1.0) < 1e-12)


def test_one_turn_map_symplectic_dba_ring():
    """|det J - 1| < 1e-12 on the one-turn map of the DA ring, over the DA
    grid corner to corner."""
    ring = lattices.dba_ring(16)
    pool = lattices.da_grid(16, 4)
    tm = ufo.TangentMap(ring, flags=ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED, turns=1, particles=len(pool),
                        parameters=["x", "y"])
    tm.parameters[:] = pool
    tm.run()
    det = np.linalg.det(tm.jacobian)
    print(f"max |det J - 1| = {np.max(np.abs(det - 1)):.2e}, max |J^T S J - S| = {symplectic_error(tm.jacobian).max():.2e}")
    assert np.all(np.abs(det - 1.0) < 1e-12)
    assert np.all(symplectic_error(tm.jacobian) < 1e-12)


def test_closed_orbit_howto_table_fp64():
    """HOWTO.md:494-503: ALBA, each of the first 8 sextupoles displaced by
    1 mm.  The documented orbits are the reference's Nelder-Mead results,
    whose own convergence error is a few 1e-12 (the oracle's restated amoeba
    reproduces them digit for digit); the Newton orbits agree within 1e-11
    and are fixed points of the oracle's one-turn map to 1e-15."""
    g = np.array(G["alba_co_sextupole_dx_fp64"]["orbits"])
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    sext = alba.RING.find(lambda e: isinstance(e, ufo.Sextupole))
    prm = [(e, "dx") for e in sext]
    co = ufo.ClosedOrbit(alba.RING, particles=len(sext), flags=ufo.FIVED | ufo.DOUBLE_PRECISION, parameters=prm)
    co.parameters[:] = np.identity(len(sext)) * 1e-3
    co.run()
    print("Newton steps:", np.bincount(co.steps))
    np.testing.assert_allclose(co.orbits[:8, :2], g[:, :2], rtol=0, atol=1e-11)
    assert np.all(np.abs(co.orbits[:8, 2:]) < 1e-13)
    # fixed point of the reference's one-turn map (oracle), all 120 orbits
    pool = np.concatenate([co.orbits, co.parameters], axis=1)
    t = runner.track(alba.RING, pool, flags=ufo.FIVED | ufo.DOUBLE_PRECISION, turns=1,
                     parameters=C4 + prm, where=[-1])[:, 0, :4]
    assert np.max(np.abs(t - co.orbits)) < 1e-15



def test_closed_orbit_default_fp32_flags():
    """tests/closed_orbit.py as the reference runs it, with the default flags
    (FP32, thick elements): ALBA with QD1 displaced by 0.1 mm in x and y.
    One FP32 Track turn from the orbit returns to it within the reference
    test's 1e-6 (here 5e-8: FP32 rounding over 768 elements; 5e-9 measured), and the orbit is within 1e-7 of the oracle's
    FP32 Nelder-Mead orbit (which lies 1.4e-8 from the FP64 orbit)."""
    lat = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    lat.QD1.dy = 1e-4
    lat.QD1.dx = 1e-4
    co = ufo.ClosedOrbit(lat.RING, particles=1)
    co.run()
    assert co.fp32
    tr = ufo.Track(lat.RING, parameters=C4, turns=1, particles=1, where=[-1])
    tr.parameters[0] = co.orbits[0]
    tr.run()
    delta = np.abs(tr.tracks[0, 0, :4].astype(np.float64) - co.orbits[0])
    print("FP32 orbit", co.orbits[0], "one-turn delta", delta, "Newton steps", co.steps[0])
    assert np.max(delta) < 5e-8
    ref = runner.closed_orbit(lat.RING, np.zeros((1, 1)), flags=0, parameters=[])
    assert np.max(np.abs(co.orbits[0] - ref[0])) < 1e-7

def test_closed_orbit_misaligned_quad_vs_tracking():
    """tests/closed_orbit.py: ALBA with QD1 displaced by 0.1 mm in x and y;
    one turn of Track from the orbit returns to it (reference tolerance 1e-6,
    here 1e-14)."""
    lat = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    lat.QD1.dy = 1e-4
    lat.QD1.dx = 1e-4
    co = ufo.ClosedOrbit(lat.RING, particles=1, flags=ufo.DOUBLE_PRECISION)
    co.run()
    tr = ufo.Track(lat.RING, parameters=C4, turns=1, particles=1, where=[-1], flags=ufo.DOUBLE_PRECISION)
    tr.parameters[0] = co.orbits[0]
    tr.run()
    assert np.max(np.abs(tr.tracks[0, 0, :4] - co.orbits[0])) < 1e-14
    assert np.max(np.abs(co.orbits[0])) > 1e-6
