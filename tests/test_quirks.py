"""The reference's quirks that the drop-in keeps (or deliberately fixes),
one test each, on the CPU: the element lowering and the emitted kernels
are compared with the oracle, which restates the reference."""
import numpy as np
import pytest

import ufo_amd as ufo
from ufo_amd import codegen, lattices
from oracle import runner

import host_emulation

COORD6 = ["x", "px", "y", "py", "dp"]
FL = ufo.KICK | ufo.DOUBLE_PRECISION


def _bunch(n=32, seed=4):
    rng = np.random.default_rng(seed)
    return np.stack([rng.normal(0, 2e-3, n), rng.normal(0, 1e-4, n), rng.normal(0, 2e-3, n),
                     rng.normal(0, 1e-4, n), rng.normal(0, 1e-3, n)], 1)


def _ring(elem):
    return ufo.Line("R", [ufo.Drift("D1", length=0.5), elem, ufo.Drift("D2", length=0.5),
                          ufo.Quadrupole("QF", length=0.3, k1=1.1), ufo.Drift("D3", length=0.5),
                          ufo.Quadrupole("QD", length=0.3, k1=-1.1)])


@pytest.mark.parametrize("cls,main,skew", [(ufo.Sextupole, "k2", "k2s"), (ufo.Octupole, "k3", "k3s")])
def test_skew_main_field_is_ignored(cls, main, skew):
    """ufo/methods.py:219,227 pass the error list dks to teapot and drop
    ksl, so k2s (sextupole) and k3s (octupole) have no effect: the lowering
    with a skew strength equals the lowering without, and the oracle agrees
    that the tracks are identical."""
    a = cls("S", length=0.2, **{main: 30.0, skew: 7.0})
    b = cls("S", length=0.2, **{main: 30.0, skew: 0.0})
    assert a.method(flags=FL) == b.method(flags=FL)
    pool = _bunch()
    ta = runner.track(_ring(a), pool, flags=FL, turns=20, parameters=COORD6, where=[-1])
    tb = runner.track(_ring(b), pool, flags=FL, turns=20, parameters=COORD6, where=[-1])
    np.testing.assert_array_equal(ta, tb)


def test_numeric_zero_override_falls_back_to_the_stored_value():
    """Element.method resolves ``value or self.value`` (ufo/elements.py:201-207):
    an explicit numeric 0 falls back to the stored strength, while a
    parameter string binds."""
    q = ufo.Quadrupole("Q", length=0.5, k1=1.2)
    assert q.method(flags=FL, k1=0) == q.method(flags=FL)
    bound = q.method(flags=FL, k1="parameters[0]")
    assert bound != q.method(flags=FL)
    assert any(e.params() == {0} for op in bound[0] for e in op.exprs())


def test_mutable_defaults_are_not_shared_and_where_is_not_mutated():
    """The reference's default lists are shared between instances and
    Optics appends -1 to the caller's ``where`` (ufo/optics.py:55-56); here
    every element owns its lists and the caller's list is left alone."""
    a, b = ufo.Quadrupole("A"), ufo.Quadrupole("B")
    a.dkn.append(1.0)
    assert b.dkn == []
    line = _ring(ufo.Marker("M"))
    where = [2]
    o = ufo.Optics(line, where=where, flags=ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION)
    assert where == [2] and o.count == 2


def test_track_without_parameters_is_accepted():
    """ufo/track.py:57 sizes the pool [particles, len(parameters)], so
    parameters=[] fails there; here the pool has one unused column and every
    particle starts at the origin (which a linear ring keeps)."""
    line = _ring(ufo.Marker("M"))
    k = codegen.track_kernel(line, FL, 5, [], [-1], 0.0)
    host = host_emulation.run_track(k, np.zeros((4, 1)), 5)
    assert np.all(host == 0.0)
    tr = ufo.Track(line, flags=FL, turns=5, particles=4, parameters=[], where=[-1])
    assert tr.parameters.shape == (4, 1)


def test_nan_start_is_lost_at_turn_zero():
    """The kernels keep IEEE semantics, so isnan(x) is never folded away
    (under -cl-fast-relaxed-math the reference's compiler may assume no
    NaN): a particle whose x is NaN stops after its first turn, exactly as
    in the strict oracle."""
    line = _ring(ufo.Marker("M"))
    pool = _bunch(4)
    pool[1, 0] = np.nan
    k = codegen.track_kernel(line, FL, 6, COORD6, [-1], 0.0)
    assert "isnan(x)" in k.src
    host = host_emulation.run_track(k, pool, 6)
    ref = runner.track(line, pool, flags=FL, turns=6, parameters=COORD6, where=[-1])
    np.testing.assert_array_equal(np.isnan(host[:, :, 0]), np.isnan(ref[:, :, 0]))
    assert np.all(np.isnan(host[1, :, 0])) and not np.any(np.isnan(host[0, :, 0]))


def test_duplicate_observation_point_loses_the_particle_as_in_the_reference():
    """where = [3, 3, -1] on a KICK ring whose element 3 is a quadrupole: the
    reference fractures the element at 0 a second time (ufo/lattice.py:101-104
    removes one 0 and passes the other to the element), and the zero-length
    teapot piece evaluates its weak-focusing term as 0/0 (ufo/methods.py), so
    px turns NaN at the duplicate point and the particle is lost in turn 0.
    The emitted kernel does the same: px is NaN in the second sample, every
    later sample is NaN and the loss flags agree with the oracle.  (The
    oracle's strict build also carries the NaN into x through the
    zero-length drift, px * 0; the kernels drop zero terms, as a compiler
    under -cl-fast-relaxed-math may, so there x turns NaN at the next drift.)"""
    line = lattices.fodo_ring(4)
    assert isinstance(line.flatten()[3], ufo.Quadrupole)
    pool = _bunch(8)
    k = codegen.track_kernel(line, FL, 2, COORD6, [3, 3, -1], 0.0)
    host = host_emulation.run_track(k, pool, 6)
    ref = runner.track(line, pool, flags=FL, turns=2, parameters=COORD6, where=[3, 3, -1])
    assert np.all(np.isnan(host[:, 1, 1])) and np.all(np.isnan(ref[:, 1, 1]))
    assert np.all(np.isnan(host[:, 2:, :4])) and np.all(np.isnan(ref[:, 2:, :4]))
    np.testing.assert_allclose(host[:, 0], ref[:, 0], rtol=1e-12, atol=1e-18)


def test_indexed_strength_attributes():
    """``dkN``/``dkNs`` read and write single entries of an element's error
    lists and ``kN``/``kNs`` those of a Multipole's strength lists
    (ufo/elements.py:37-50,128-145); ``dkN`` keyword overrides of method()
    extend the error list with zeros to reach N (ufo/elements.py:14-23),
    ``kN`` overrides of a Multipole need the entry to exist."""
    q = ufo.Quadrupole("Q", length=0.5, k1=1.0, dkn=[0.0, 0.0, 3.0], dks=[0.0, 1.5])
    assert (q.dk2, q.dk1s) == (3.0, 1.5)
    q.dk2 = 4.0
    q.dk0s = -2.0
    assert q.dkn == [0.0, 0.0, 4.0] and q.dks == [-2.0, 1.5]
    for missing in ("dk7", "dk5s", "dkx", "dk"):
        with pytest.raises(AttributeError):
            getattr(q, missing)
    m = ufo.Multipole("M", knl=[0.0, 0.5, 10.0], ksl=[0.0, 0.25])
    assert (m.k1, m.k2, m.k1s) == (0.5, 10.0, 0.25)
    m.k2 = 12.0
    m.k0s = 0.125
    assert m.knl == [0.0, 0.5, 12.0] and m.ksl == [0.125, 0.25]
    # a dkN override beyond the list pads with zeros; the same strength given
    # through dkn= lowers to the same ops
    grown = ufo.Sextupole("S", length=0.2, k2=30.0).method(flags=FL, dk4=7.0, dk3s=2.0)
    given = ufo.Sextupole("S", length=0.2, k2=30.0).method(flags=FL, dkn=[0.0, 0.0, 0.0, 0.0, 7.0],
                                                           dks=[0.0, 0.0, 0.0, 2.0])
    assert grown == given
    assert ufo.Multipole("M", knl=[0.0, 0.5]).method(flags=FL, k1=0.75) == \
        ufo.Multipole("M", knl=[0.0, 0.75]).method(flags=FL)
    with pytest.raises(IndexError):
        ufo.Multipole("M", knl=[0.0, 0.5]).method(flags=FL, k3=1.0)


def test_constructor_signatures_bind_like_the_reference():
    """Each element's constructor takes the reference's arguments in the
    reference's order (ufo/elements.py:74,93,124,189,244,310,380,435,487,512):
    positional and keyword calls bind the same way, an unknown keyword or a
    surplus positional raises TypeError, and ``inspect.signature`` shows the
    arguments with their defaults."""
    import inspect
    order = {ufo.Drift: ["label", "length"],
             ufo.Multipole: ["label", "knl", "ksl", "dx", "dy"],
             ufo.Quadrupole: ["label", "slices", "length", "k1", "dx", "dy", "dkn", "dks"],
             ufo.Sbend: ["label", "slices", "length", "angle", "k1", "e1", "e2", "hgap", "fint", "dx", "dy",
                         "dkn", "dks"],
             ufo.Rbend: ["label", "slices", "length", "angle", "k1", "e1", "e2", "hgap", "fint", "dx", "dy",
                         "dkn", "dks"],
             ufo.Sextupole: ["label", "length", "slices", "k2", "k2s", "dx", "dy", "dkn", "dks"],
             ufo.Octupole: ["label", "length", "slices", "k3", "k3s", "dx", "dy", "dkn", "dks"],
             ufo.Wire: ["label", "x", "y", "k"],
             ufo.Aperture: ["label", "window"],
             ufo.Marker: ["label"]}
    for cls, names in order.items():
        assert list(inspect.signature(cls).parameters) == names, cls
    q = ufo.Quadrupole("Q", 6, 0.5, 1.25, 1e-4)
    assert (q.slices, q.length, q.k1, q.dx, q.dy, q.dkn) == (6, 0.5, 1.25, 1e-4, 0.0, [])
    s = ufo.Sextupole("S", 0.2, None, 30.0)
    assert (s.length, s.slices, s.k2) == (0.2, ufo.DEFAULT_SEXTUPOLE_SLICES, 30.0)
    with pytest.raises(TypeError):
        ufo.Drift("D", length=1.0, k1=2.0)
    with pytest.raises(TypeError):
        ufo.Drift("D", 1.0, 2.0)
    with pytest.raises(TypeError):
        ufo.Quadrupole()
    # list arguments are copied, and the slice default is read when the
    # element is built
    errors = [0.0, 1e-4]
    q = ufo.Quadrupole("Q", length=0.5, dkn=errors)
    errors.append(5.0)
    assert q.dkn == [0.0, 1e-4]
    saved = ufo.DEFAULT_QUADRUPOLE_SLICES
    try:
        ufo.DEFAULT_QUADRUPOLE_SLICES = 3
        assert ufo.Quadrupole("Q", length=0.5).slices == 3
    finally:
        ufo.DEFAULT_QUADRUPOLE_SLICES = saved
    # repr walks the reference's parameters list; Rbend's integer defaults
    assert repr(ufo.Drift("D", 1.5)) == "D       : DRIFT        length=1.5"
    assert "hgap=0   fint=0" in repr(ufo.Rbend("R", length=1.0, angle=0.1))
    assert "hgap=0.0   fint=0.0" in repr(ufo.Sbend("B", length=1.0, angle=0.1))
