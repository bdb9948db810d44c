"""The dump formats byte for byte (ufo/lattice.py:126-202 and the element
``dump`` methods of ufo/elements.py): a lattice holding every element kind
written as mad, elegant, at and opa, compared with the text the reference's
format strings give.  Where the reference's writer fails, the drop-in
differs, and each such divergence has a test of exactly that difference:

* Multipole in mad/elegant: the reference writes ``KNL=`` twice, the second
  holding the skew list (ufo/elements.py:173); here the second is ``KSL=``.
* Sbend in at/opa, Wire and Aperture in every format: the reference prints
  a message and returns None, on which its lattice ``dump`` raises
  TypeError; here the message is printed and nothing is written.
* Rbend in opa: the reference prints a message and ends the interpreter
  (``exit()``, ufo/elements.py:361-362); here the message is printed and
  nothing is written.
* Sextupole and Octupole in at: the reference's format string names an
  undefined ``label`` (NameError, ufo/elements.py:419,474); here the
  element's label is written.
* A Line in an unknown style: the reference fails on an unbound local;
  here ValueError.
"""

import pytest

import ufo_amd as ufo


def _lattice(kinds):
    lat = ufo.Lattice()
    every = {
        "M": ufo.Marker("M"),
        "D": ufo.Drift("D", length=1.5),
        "Q": ufo.Quadrupole("Q", length=0.5, k1=1.25),
        "B": ufo.Sbend("B", length=1.25, angle=0.125, k1=-0.5, e1=0.0625, e2=0.03125),
        "R": ufo.Rbend("R", length=1.0, angle=0.25),
        "S": ufo.Sextupole("S", length=0.25, k2=12.5),
        "O": ufo.Octupole("O", length=0.5, k3=100.0),
        "P": ufo.Multipole("P", knl=[0.0, 0.5, 4.0], ksl=[0.0, 0.25]),
    }
    for k in kinds:
        setattr(lat, k, every[k])
    lat.L = ufo.Line("L", [every[k] for k in kinds])
    return lat


def _dump(lat, style, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    ufo.dump(lat, f"lat.{style}", style=style)
    with open(f"lat.{style}", newline="") as f:
        return f.read()


MAD_ELEMENTS = ("M: MARKER;\n"
                "D: DRIFT, L=1.5;\n"
                "Q: QUADRUPOLE, L=0.5, K1=1.25;\n"
                "B: SBEND, L=1.25, ANGLE=0.125, K1=-0.5, E1=0.0625, E2=0.03125;\n"
                "R: RBEND, L=1.0, ANGLE=0.25, K1=0.0;\n"
                "S: SEXTUPOLE, L=0.25, K2=12.5;\n"
                "O: OCTUPOLE, L=0.5, K3=100.0;\n")


@pytest.mark.parametrize("style,line_end", [("mad", ");\n\n"), ("elegant", ")\n\n")])
def test_mad_and_elegant(tmp_path, monkeypatch, style, line_end):
    text = _dump(_lattice("MDQBRSOP"), style, tmp_path, monkeypatch)
    assert text == (MAD_ELEMENTS
                    + "P: MULTIPOLE, KNL={ 0.0, 0.5, 4.0 }, KSL={ 0.0, 0.25 };\n"
                    + "L: LINE=(M, D, Q, B, R, S, O, P" + line_end
                    + "\n")


def test_multipole_skew_list_is_written_as_ksl():
    """The divergence itself: the reference's line would read
    ``KNL={ 0.0, 0.5, 4.0 }, KNL={ 0.0, 0.25 }``."""
    line = ufo.Multipole("P", knl=[0.0, 0.5, 4.0], ksl=[0.0, 0.25]).dump("mad")
    assert line == "P: MULTIPOLE, KNL={ 0.0, 0.5, 4.0 }, KSL={ 0.0, 0.25 };\n"
    assert line.replace("KSL=", "KNL=") == "P: MULTIPOLE, KNL={ 0.0, 0.5, 4.0 }, KNL={ 0.0, 0.25 };\n"


def test_at(tmp_path, monkeypatch):
    text = _dump(_lattice("MDQRSO"), "at", tmp_path, monkeypatch)
    assert text == ("function lat\n\n"
                    "global FAMLIST THERING GLOBVAL;\n"
                    "E_0 = 510998.94999999995;\n"
                    "GLOBVAL.E0 = 3000000000.0-E_0;\n"
                    "GLOBVAL.LatticeFile = 'lat';\n"
                    "FAMLIST = cell(0);\n"
                    "THERING = cell(0);\n\n"
                    "M = marker('M', 'IdentityPass');\n"
                    "D = drift('D', 1.5, 'DriftPass');\n"
                    "Q = quadrupole('Q', 0.5, 1.25,'StrMPoleSymplectic4Pass');\n"
                    "R = rbend('R', 1.0, 0.25, 0, 0, 0.0, 'BndMPoleSymplectic4Pass');\n"
                    "S = sextupole('S', 0.25, 12.5,'StrMPoleSymplectic4Pass');\n"
                    "O = octupole('O', 0.5, 100.0,'StrMPoleSymplectic4Pass');\n"
                    "L = [M, D, Q, R, S, O];\n\n"
                    "\n"
                    "\n\nbuildlat(RING);\n"
                    "% Set all magnets to same energy\n"
                    "THERING = setcellstruct(THERING, 'Energy', 1:length(THERING), GLOBVAL.E0);\n"
                    "evalin('caller', 'global THERING FAMLIST GLOBVAL');\n"
                    "if nargout\n    varargout{1} = THERING;\nend\nend")


def test_opa(tmp_path, monkeypatch):
    text = _dump(_lattice("MDQSO"), "opa", tmp_path, monkeypatch)
    assert text == ("energy = 3.0;\r\n"
                    "M : opticsmarker;\r\n"
                    "D : drift, l = 1.5;\r\n"
                    "Q : quadrupole, l = 0.5, k = 1.25;\r\n"
                    "S : sextupole, l = 0.25, k = 6.25;\r\n"
                    "O : octupole, l = 0.5, k = 50.0;\r\n"
                    "L : M, D, Q, S, O;\r\n"
                    "\r\n")


def test_sextupole_and_octupole_at_lines_carry_the_label():
    """The divergence itself: the reference's format string raises
    NameError here."""
    assert ufo.Sextupole("S", length=0.25, k2=12.5).dump("at").startswith("S = sextupole('S', ")
    assert ufo.Octupole("O", length=0.5, k3=100.0).dump("at").startswith("O = octupole('O', ")


def test_unsupported_combinations_write_nothing(capsys, tmp_path, monkeypatch):
    """Sbend in at/opa, Rbend in opa, Wire and Aperture anywhere: a message
    and no text (the reference returns None, or exits for Rbend/opa), so a
    lattice holding them still dumps the rest."""
    b = ufo.Sbend("B", length=1.0, angle=0.1)
    assert b.dump("at") == "" and b.dump("opa") == ""
    assert "NOT SUPPORTED" in capsys.readouterr().out
    assert ufo.Rbend("R", length=1.0, angle=0.1).dump("opa") == ""
    assert "RBEND" in capsys.readouterr().out
    for e in (ufo.Wire("W", x=0.01, k=1e-6), ufo.Aperture("A", window="(fabs(x) < 0.1)")):
        for style in ("mad", "elegant", "at", "opa"):
            assert e.dump(style) == ""
    assert "not supported" in capsys.readouterr().out
    lat = _lattice("MD")
    lat.B = b
    text = _dump(lat, "opa", tmp_path, monkeypatch)
    assert text == "energy = 3.0;\r\nM : opticsmarker;\r\nD : drift, l = 1.5;\r\nL : M, D;\r\n\r\n"


def test_multipole_in_at_or_opa_raises():
    p = ufo.Multipole("P", knl=[0.0, 1.0])
    for style in ("at", "opa"):
        with pytest.raises(Exception, match="not supported"):
            p.dump(style)


def test_line_in_unknown_style_raises():
    with pytest.raises(ValueError):
        ufo.Line("L", [ufo.Drift("D", length=1.0)]).dump("madx")
