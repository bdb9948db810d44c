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


def test_reader_values_of_a_golden_file():
    """Every value of tests/golden/lattices/elements.mad (a copy of the
    reference's optics/elements.mad) lands on the attribute the reference's
    constructor table names."""
    lat = ufo.Lattice(os.path.join(OPTICS_DIR, "elements.mad"))
    assert (lat.Q.length, lat.Q.k1) == (0.8786, 1.5878)
    assert (lat.B.length, lat.B.angle, lat.B.k1, lat.B.e1, lat.B.e2) == (0.4929, 0.234, -2.4987, 0.1, -0.1)
    assert lat.D.length == 0.2876 and isinstance(lat.M, ufo.Marker)
    assert (lat.S.length, lat.S.k2, lat.S.k2s) == (1.8776, 167.44, 0.0)
    assert lat.O.knl == [0.0, 0.0, 0.0, 8765.7] and lat.O.ksl == []
    assert [e.label for e in lat.TEST] == ["Q", "B", "D", "M", "S", "O"]
    assert list(vars(lat)) == ["Q", "B", "D", "M", "S", "O", "TEST"]


def test_reader_dialect_rules(capsys):
    from ufo_amd import mad
    t = mad.loads('Q_1: quadrupole, L=1., K1=-2;\n'
                  'W: APERTURE, WINDOW="(fabs(x) < 0.1)";\n'
                  'V: MULTIPOLE, KNL={ 0, -0.5 }, KSL={};\n'
                  'X: KICKER, L=2;\r\n'                       # unknown kind: skipped
                  'S: SEXTUPOLE ( L=0.2 , K2=3;\n'             # one stray token in place of a comma
                  'R: LINE=(Q_1,W , V, S);\n'
                  'RR: LINE = (R, Q_1);')
    out = capsys.readouterr().out
    assert "unknown statement 'KICKER'" in out and "expected ','" in out
    assert (t["Q_1"].length, t["Q_1"].k1) == (1.0, -2.0)
    assert t["W"].window == "(fabs(x) < 0.1)"
    assert t["V"].knl == [0.0, -0.5] and t["V"].ksl == []
    assert (t["S"].length, t["S"].k2) == (0.2, 3.0)
    assert "X" not in t and [e.label for e in t["RR"].flatten()] == ["Q_1", "W", "V", "S", "Q_1"]
    assert t["RR"][0] is t["R"]
    for bad, err in [("Q: QUADRUPOLE, L=1e-3;", RuntimeError),            # no exponent in the dialect
                     ("Q: QUADRUPOLE, L=.5;", RuntimeError),              # no leading '.'
                     ("Q: QUADRUPOLE, L=1, K1=2 K2=3;", RuntimeError),    # a missing comma before a key
                     ("Q: QUADRUPOLE, l=1;", TypeError),                  # keys are uppercase
                     ("Q QUADRUPOLE, L=1;", RuntimeError),
                     ("R: LINE=(Q);", KeyError),                          # undefined member
                     ("R: LINE=();", RuntimeError),
                     ("Q: DRIFT, L=1; ! comment", RuntimeError),
                     ("Q: DRIFT, L=1", RuntimeError)]:
        with pytest.raises(err):
            mad.loads(bad)
    with pytest.raises(RuntimeError, match="line 3"):
        mad.loads("A: DRIFT, L=1;\nB: DRIFT, L=2;\nC: DRIFT, L=#;")


def test_multipole_dump_reads_back(tmp_path):
    """The mad form of a Multipole (blanks inside the braces) loads back
    with both strength lists."""
    lat = ufo.Lattice()
    lat.M = ufo.Multipole("M", knl=[0.0, 0.5, 12.25], ksl=[0.0, -0.25])
    lat.D = ufo.Drift("D", length=1.5)
    lat.L = ufo.Line("L", [lat.M, lat.D])
    path = str(tmp_path / "m.mad")
    ufo.dump(lat, path, style="mad")
    again = ufo.Lattice(path)
    assert again.M.knl == [0.0, 0.5, 12.25] and again.M.ksl == [0.0, -0.25] and again.D.length == 1.5
