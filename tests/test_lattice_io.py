"""Lattice files and geometry (ufo/mad.py, ufo/lattice.py:54-84,126-202):
the ALBA lattice's element counts, length and bending angle, ``find`` /
``locate`` / ``survey``, and the ``dump`` formats with a mad round trip that
keeps the optics."""
import collections
import os

import numpy as np
import pytest

import ufo_amd as ufo
from oracle import runner

from conftest import OPTICS_DIR


def _alba():
    return ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))


def test_alba_counts_length_angle():
    ring = _alba().RING
    flat = ring.flatten()
    kinds = collections.Counter(type(e).__name__ for e in flat)
    assert len(flat) == 768
    assert kinds == {"Drift": 504, "Sextupole": 120, "Quadrupole": 112, "Sbend": 32}
    assert ring.length == pytest.approx(268.800296, abs=1e-6)
    assert ring.angle == pytest.approx(2 * np.pi, abs=2e-9)


@pytest.mark.parametrize("name", ["alba", "align", "bend", "drift", "elements", "fodo", "fodo2", "multipole",
                                  "oct", "triplet"])
def test_every_golden_lattice_loads(name):
    lat = ufo.Lattice(os.path.join(OPTICS_DIR, f"{name}.mad"))
    lines = [v for v in vars(lat).values() if isinstance(v, ufo.Line)]
    assert lines and all(len(l.flatten()) > 0 for l in lines)


def test_find_and_locate():
    ring = _alba().RING
    flat = ring.flatten()
    quads = ring.find(lambda e: isinstance(e, ufo.Quadrupole))
    assert len(quads) == 112 and all(isinstance(flat[i], ufo.Quadrupole) for i in quads)
    i = quads[3]
    s = sum(getattr(e, "length", 0.0) for e in flat[:i])
    assert ring.locate(i) == pytest.approx(s)
    assert ring.locate(i + 0.5) == pytest.approx(s + 0.5 * flat[i].length)
    assert ring.locate(-1) == pytest.approx(ring.length)


def test_survey_closes_the_ring():
    """The survey of the flattened ring returns to its start: 32 bends of
    the ALBA ring turn by 2 pi (to the rounding of the angles in the file)."""
    flat = ufo.Line("F", _alba().RING.flatten())
    xs, alphas = flat.survey()
    assert alphas[-1] == pytest.approx(2 * np.pi, abs=2e-9)
    assert np.hypot(*xs[-1]) < 1e-6
    assert max(np.hypot(x, y) for x, y in xs) > 50.0


def test_dump_mad_round_trip_keeps_the_optics(tmp_path):
    """Everything the mad format carries survives dump + load.  The SBEND
    line has no HGAP/FINT (as the reference writes it, ufo/elements.py
    Sbend.dump), so the fringe terms are set to zero before the comparison."""
    alba = _alba()
    for e in alba.RING.flatten():
        if isinstance(e, ufo.Sbend):
            e.hgap = e.fint = 0.0
    path = str(tmp_path / "alba_dump.mad")
    ufo.dump(alba, path, style="mad")
    again = ufo.Lattice(path)
    a, b = alba.RING.flatten(), again.RING.flatten()
    assert [type(e).__name__ for e in a] == [type(e).__name__ for e in b]
    for e, f in zip(a, b):
        for attr in ("length", "k1", "k2", "angle", "e1", "e2"):
            if hasattr(e, attr):
                assert getattr(f, attr) == pytest.approx(getattr(e, attr), rel=1e-12, abs=1e-15)
    fl = ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION
    q = []
    for lat in (alba, again):
        o = runner.Optics(lat.RING, where=[], flags=fl)
        o.run()
        q.append((o.qx, o.qy))
    assert q[0] == pytest.approx(q[1], abs=1e-12)


@pytest.mark.parametrize("style", ["elegant", "at", "opa"])
def test_dump_other_formats(tmp_path, style):
    alba = _alba()
    path = str(tmp_path / f"alba.{style}")
    ufo.dump(alba, path, style=style)
    text = open(path).read()
    assert len(text) > 1000
    assert text.count("QD1") > 0 or text.count("qd1") > 0
