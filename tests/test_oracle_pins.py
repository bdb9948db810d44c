"""Pins of the oracle (oracle/refgen.py + oracle/runner.py) against the
numbers the reference's own tests and docs hold, before the oracle is used
to check the HIP kernels.  CPU only."""
import json
import os

import numpy as np
import pytest

import ufo_amd as ufo
from oracle import runner

from conftest import OPTICS_DIR, ROOT

G = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_values.json")))


def lattice(name):
    return ufo.Lattice(os.path.join(OPTICS_DIR, f"{name}.mad"))


@pytest.mark.parametrize("double", [False, True])
@pytest.mark.parametrize("mode", ["strict", "fast"])
def test_tracking_madx_vector(double, mode):
    """tests/tracking.py: thick FODO, FIVED, x0 = 1 mm, 10 turns vs MAD-X
    (reference tolerance 1e-6; the maps reproduce it to ~5e-9)."""
    g = G["tracking_fodo_x_madx"]
    flags = ufo.FIVED | (ufo.DOUBLE_PRECISION if double else 0)
    pool = np.full((2, 1), g["x0"])
    t = runner.track(lattice("fodo").RING, pool, flags=flags, turns=10, parameters=["x"], where=[-1], mode=mode)
    for p in range(2):
        assert np.max(np.abs(t[p, :, 0] - np.array(g["x"]))) < 1e-8


def test_aperture_window_is_keep_condition():
    """Aperture splices ``if (<window> == 0) x = NAN;`` (ufo/elements.py:518).
    With C precedence ``x > 0.1 == 0`` is ``(x > 0.1) == 0``: the window is
    the keep condition, as documented at ufo/elements.py:509-510 — x = 0.2 is
    kept and x = 0 is lost.  (tests/aperture.py expects the opposite; under
    C precedence that expectation fails on the reference itself.)"""
    line = ufo.Line("a_line", [ufo.Aperture("ap", window="x > 0.1")])
    t = runner.track(line, np.array([[0.0], [0.2]]), flags=ufo.FIVED, turns=1, parameters=["x"], where=[-1])
    assert np.isnan(t[0, 0, 0])
    assert t[1, 0, 0] == np.float32(0.2)


def test_aperture_precedence_splice():
    """``a && b == 0`` binds as ``a && (b == 0)``: window text is spliced, not
    parenthesised."""
    line = ufo.Line("l", [ufo.Aperture("ap", window="x > 0.1 && y > 0.1")])
    pool = np.array([[0.2, 0.0], [0.2, 0.2], [0.0, 0.0]])
    t = runner.track(line, pool, flags=ufo.FIVED | ufo.DOUBLE_PRECISION, turns=1, parameters=["x", "y"], where=[-1])
    # x>0.1 && (y>0.1)==0 -> lost when x > 0.1 and y <= 0.1
    assert np.isnan(t[0, 0, 0]) and t[1, 0, 0] == 0.2 and t[2, 0, 0] == 0.0


def test_alba_tunes_madx():
    """tests/optics.py: ALBA periodic tunes vs MAD-X within 1e-5 (FP32 as in
    the reference test); FP64 agrees to 1e-8."""
    g = G["alba_tunes_madx"]
    o = runner.Optics(lattice("alba").RING, where=[])
    o.run()
    assert abs(o.qx - g["qx"]) < 1e-5 and abs(o.qy - g["qy"]) < 1e-5
    o = runner.Optics(lattice("alba").RING, where=[], flags=ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION)
    o.run()
    assert abs(o.qx - g["qx"]) < 1e-8 and abs(o.qy - g["qy"]) < 1e-8


def test_alba_chromaticity_fd_madx():
    """tests/optics3.py: tune shift with dp (step 1e-4), LINEAR, vs MAD-X
    within 0.1."""
    g = G["alba_chromaticity_fd_madx"]
    o = runner.Optics(lattice("alba").RING, where=[], parameters=["dp"], flags=ufo.LINEAR | ufo.DOUBLE_PRECISION)
    o.parameters[0] = 0.0
    o.run()
    qx0, qy0 = o.qx, o.qy
    o.parameters[0] = 1e-4
    o.run()
    assert abs((o.qx - qx0) / 1e-4 - g["dqx"]) < 0.1
    assert abs((o.qy - qy0) / 1e-4 - g["dqy"]) < 0.1


def test_thick_vs_kick_tunes():
    """tests/thick_vs_kick.py: teapot (16 quad / 48 bend slices) and thick
    tunes agree within 5e-4."""
    alba = lattice("alba")
    for e in alba.RING.flatten():
        if isinstance(e, ufo.Quadrupole):
            e.slices = 16
        if isinstance(e, ufo.Sbend):
            e.slices = 48
    fl = ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION
    kick = runner.Optics(alba.RING, flags=fl | ufo.KICK)
    thick = runner.Optics(alba.RING, flags=fl)
    kick.run()
    thick.run()
    assert abs(kick.qx - thick.qx) < 5e-4 and abs(kick.qy - thick.qy) < 5e-4


def test_align_equivalent_bend():
    """tests/align.py: a quad displaced by dx=1e-4 gives the dispersion of a
    combined bend with angle 1e-4 (within 1e-6)."""
    lat = lattice("align")
    lat.QF.dx = 1e-4
    a = runner.Optics(lat.FODO, where=[-1])
    b = runner.Optics(lat.EQUIVALENT, where=[-1])
    a.run()
    b.run()
    assert abs(a.dx[0] - b.dx[0]) < 1e-6


def test_closed_orbit_howto_table_fp64():
    """HOWTO.md:494-503: ALBA, each sextupole displaced by 1 mm in turn,
    Nelder-Mead closed orbit in FP64.  The restated amoeba and thick maps
    reproduce every printed digit (9 significant) of the first 8 rows,
    including the ~1e-14 vertical residuals."""
    g = np.array(G["alba_co_sextupole_dx_fp64"]["orbits"])
    alba = lattice("alba")
    sext = alba.RING.find(lambda e: isinstance(e, ufo.Sextupole))
    prm = [(e, "dx") for e in sext]
    pool = (np.identity(len(sext)) * 1e-3)[:8]
    orb = runner.closed_orbit(alba.RING, pool, flags=ufo.FIVED | ufo.DOUBLE_PRECISION, parameters=prm)
    np.testing.assert_allclose(orb, g, rtol=6e-9, atol=0)


def test_alba_sa_grid_howto_fp32():
    """HOWTO.md:332-370: 12x12 ALBA grid, FP32, 1000 turns.  With the fast
    build every cell whose documented value is 0 or 1000 is reproduced; the
    grid is mirror-symmetric in y; only boundary cells may differ."""
    g = G["alba_sa_grid_fp32"]
    n = g["count"]
    ref = np.array(g["lost"])
    x = np.linspace(-0.04, 0.04, n)
    xx, yy = np.meshgrid(x, x)
    pool = np.stack([xx.ravel(), yy.ravel()], 1)
    lost = runner.stable_aperture(lattice("alba").RING, pool, flags=ufo.FIVED, turns=g["turns"],
                                  parameters=["x", "y"], mode="fast").reshape(n, n)
    robust = (ref == 0) | (ref == 1000)
    assert np.all(lost[robust] == ref[robust])
    assert np.all(lost == lost[::-1])
    assert np.sum(lost != ref) <= 6


def test_fast_oracle_leaves_subnormals_alone():
    """Loading a fast-math oracle build does not switch the process to
    flush-to-zero: strict runs and numpy after it still see subnormals
    (gcc's crtfastmath.o, added to a -ffast-math link, would set FTZ/DAZ in
    its constructor)."""
    import ufo_amd as ufo
    from ufo_amd import lattices
    ring = lattices.fodo_ring(2)
    pool = lattices.gaussian_bunch(4)
    for mode in ("fast", "bench_fast"):
        runner.stable_aperture(ring, pool, flags=ufo.KICK | ufo.DOUBLE_PRECISION, turns=2,
                               parameters=["x", "px", "y", "py", "dp"], mode=mode)
    tiny = np.array([5e-324, 1e-310])
    assert np.array_equal(tiny * 1.0, tiny)
    assert np.nextafter(0.0, 1.0) * 3 == 1.5e-323
