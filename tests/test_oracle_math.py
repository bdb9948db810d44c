"""Pins of the oracle against what the mathematics fixes: closed forms
evaluated in mpmath, an independent integration of the bend equations,
exact symmetries and a third-party phase advance.  Each is chosen so that a
plausible slip in the restated templates (a dropped term, a wrong sign or
index, a missing factorial, a wrong slice length) fails it.  CPU only."""
import math

import mpmath as mp
import numpy as np
import pytest

import ufo_amd as ufo
from ufo_amd import lattices
from oracle import runner

mp.mp.dps = 40
F64 = ufo.DOUBLE_PRECISION
C6 = ["x", "px", "y", "py", "dp"]


def one_pass(line, pool, flags, parameters=C6, mode="strict"):
    return runner.track(line, np.atleast_2d(pool), flags=flags, turns=1, parameters=parameters,
                        where=[-1], mode=mode)[:, 0, :]


def test_multipole_kick_series_mpmath():
    """Thin multipole: px -= Re w, py += Im w, w = sum (kn + i ks) z^n / n!,
    every coefficient times 1/(1+dp) (ufo/methods.py:45-76)."""
    knl = [1.3e-4, -0.21, 3.7, -55.0, 910.0, -2.2e4, 5.5e5]
    ksl = [-2.0e-4, 0.13, -1.9, 28.0, -640.0, 1.7e4, -3.1e5]
    line = ufo.Line("m", [ufo.Multipole("M", knl=knl, ksl=ksl)])
    rng = np.random.default_rng(11)
    pts = np.stack([rng.normal(0, 3e-3, 32), rng.normal(0, 1e-4, 32), rng.normal(0, 3e-3, 32),
                    rng.normal(0, 1e-4, 32), rng.normal(0, 2e-3, 32)], 1)
    out = one_pass(line, pts, F64)
    for p, o in zip(pts, out):
        z = mp.mpc(p[0], p[2])
        w = sum(mp.mpc(knl[n], ksl[n]) * z ** n / mp.factorial(n) for n in range(len(knl)))
        w = w / (1 + mp.mpf(p[4]))
        # rounding scale: the sum of the term magnitudes (the series cancels)
        scale = float(sum(abs(mp.mpc(knl[n], ksl[n])) * abs(z) ** n / mp.factorial(n) for n in range(len(knl))))
        tol = 1e-15 * scale + 2e-16 * max(abs(p[1]), abs(p[3]))
        assert abs(o[1] - float(p[1] - w.real)) <= tol
        assert abs(o[3] - float(p[3] + w.imag)) <= tol
        assert o[0] == p[0] and o[2] == p[2]


def test_sextupole_closed_form_and_linear_truncation():
    """dpx = -k2L/2 (x^2 - y^2)/(1+dp), dpy = k2L x y/(1+dp); LINEAR drops
    orders above 1."""
    k2l = 7.5
    line = ufo.Line("s", [ufo.Multipole("S", knl=[0.0, 0.0, k2l])])
    p = np.array([2e-3, 1e-5, -1.5e-3, 2e-5, 1e-3])
    o = one_pass(line, p, F64)[0]
    x, y, d = mp.mpf(p[0]), mp.mpf(p[2]), 1 + mp.mpf(p[4])
    assert o[1] == pytest.approx(float(p[1] - k2l / 2 * (x * x - y * y) / d), rel=1e-14)
    assert o[3] == pytest.approx(float(p[3] + k2l * x * y / d), rel=1e-14)
    o = one_pass(line, p, F64 | ufo.LINEAR)[0]
    assert o[1] == p[1] and o[3] == p[3]


def test_exact_drift_closed_form():
    L = 1.7
    line = ufo.Line("d", [ufo.Drift("D", length=L)])
    p = np.array([1e-3, 3e-2, -2e-3, -4e-2, 5e-3])
    o = one_pass(line, p, F64 | ufo.EXACT)[0]
    pz = mp.sqrt((1 + mp.mpf(p[4])) ** 2 - mp.mpf(p[1]) ** 2 - mp.mpf(p[3]) ** 2)
    assert o[0] == pytest.approx(float(p[0] + p[1] * L / pz), rel=1e-15)
    assert o[2] == pytest.approx(float(p[2] + p[3] * L / pz), rel=1e-15)
    o = one_pass(line, p, F64)[0]      # paraxial: no dp dependence
    assert o[0] == pytest.approx(p[0] + p[1] * L, rel=1e-15)


def test_wire_closed_form():
    """Kick of a wire at (a, b): k (a-x, b-y) / r^2 added to (px, py)."""
    a, b, k = 0.02, -0.01, 3e-6
    line = ufo.Line("w", [ufo.Wire("W", x=a, y=b, k=k)])
    p = np.array([1e-3, 0.0, 2e-3, 0.0, 0.0])
    o = one_pass(line, p, F64)[0]
    dx, dy = mp.mpf(a) - mp.mpf(p[0]), mp.mpf(b) - mp.mpf(p[2])
    r2 = dx * dx + dy * dy
    assert o[1] == pytest.approx(float(k * dx / r2), rel=1e-14)
    assert o[3] == pytest.approx(float(k * dy / r2), rel=1e-14)


def _matrix_teapot(L, k, n):
    """2x2 teapot map of one plane in mpmath, from the slicing definition:
    end drifts L/(2(1+n)), n-1 inner drifts, n kicks of k L / n."""
    L, k = mp.mpf(L), mp.mpf(k)
    o = L / (2 * (1 + n)) if n > 1 else L / 2
    i = (L - 2 * o) / (n - 1) if n > 1 else 0
    D = lambda l: mp.matrix([[1, l], [0, 1]])
    K = mp.matrix([[1, 0], [-k * L / n, 1]])
    m = D(o)
    for s in range(n):
        m = D(o if s == n - 1 else i) * K * m
    return m


def test_linear_fodo_n_turns_matches_matrix_power():
    """Linear teapot FODO, 500 turns: the oracle equals M^500 x0 computed in
    mpmath to 1e-13 (relative to the orbit amplitude)."""
    qf = ufo.Quadrupole("QF", length=0.5, k1=1.2, slices=4)
    qd = ufo.Quadrupole("QD", length=0.5, k1=-1.2, slices=4)
    d = ufo.Drift("D", length=2.0)
    ring = ufo.Line("R", [ufo.Line("C", [qf, d, qd, d])] * 4)
    turns = 500
    x0 = np.array([[1e-3, -2e-5, 5e-4, 3e-5]])
    t = runner.track(ring, x0, flags=F64 | ufo.KICK | ufo.FIVED, turns=turns, parameters=["x", "px", "y", "py"],
                     where=[-1])
    for plane, (u, pu), sgn in (("x", (0, 1), 1), ("y", (2, 3), -1)):
        D = mp.matrix([[1, 2], [0, 1]])
        cell = D * _matrix_teapot(0.5, -sgn * 1.2, 4) * D * _matrix_teapot(0.5, sgn * 1.2, 4)
        m = (cell ** 4) ** turns
        v = m * mp.matrix([[x0[0, u]], [x0[0, pu]]])
        amp = max(abs(x0[0, u]), 1e-3)
        assert abs(t[0, -1, u] - float(v[0])) < 1e-13 * amp * 10
        assert abs(t[0, -1, pu] - float(v[1])) < 1e-13 * amp


def test_fodo_cell_phase_advance():
    """Config FODO cell (QF/QD L=0.5 k1=+-1.2, 4 slices, 2 m drifts):
    88.505 deg per cell, so 16 cells give Q = 3.933561 (computed
    independently with 2x2 matrices); the oracle's Optics reproduces the
    fractional tune."""
    ring = ufo.Line("R", [lattices.fodo_cell()] * 16)      # no aperture: the rays have x = 1
    o = runner.Optics(ring, where=[], flags=ufo.LINEAR | ufo.ACHROMATIC | F64 | ufo.KICK)
    o.run()
    assert o.qx == pytest.approx(0.933561, abs=2e-6)
    assert o.qy == pytest.approx(0.933561, abs=2e-6)


def test_thick_sbend_against_ode_integration():
    """Combined-function sector bend, thick map with dispersion
    (ufo/methods.py:142-194): compared with an RK4 integration of
    x'' = -(k1 + h^2) x + h dp, y'' = k1 y (h = angle/L, ACHROMATIC)."""
    L, ang, k1 = 1.384, 2 * math.pi / 32, -0.4
    line = ufo.Line("b", [ufo.Sbend("B", length=L, angle=ang, k1=k1)])
    p = np.array([1e-3, -2e-4, 5e-4, 1e-4, 2e-3])
    o = one_pass(line, p, F64 | ufo.ACHROMATIC)[0]
    h = ang / L
    kx = k1 + h * h

    def f(s, v):
        return np.array([v[1], -kx * v[0] + h * p[4], v[3], k1 * v[2]])

    v = np.array([p[0], p[1], p[2], p[3]])
    n = 4000
    ds = L / n
    for i in range(n):
        k_1 = f(0, v)
        k_2 = f(0, v + ds / 2 * k_1)
        k_3 = f(0, v + ds / 2 * k_2)
        k_4 = f(0, v + ds * k_3)
        v = v + ds / 6 * (k_1 + 2 * k_2 + 2 * k_3 + k_4)
    np.testing.assert_allclose(o[:4], v, rtol=1e-10, atol=1e-14)


def test_normal_multipoles_mirror_symmetry_exact():
    """A lattice of normal multipoles maps (x, px, -y, -py) to the exact
    mirror image (bitwise) — a sign error in the kick's imaginary part
    breaks it."""
    ring = lattices.fodo_ring(4)
    rng = np.random.default_rng(5)
    pool = lattices.gaussian_bunch(16, seed=2)
    mirror = pool.copy()
    mirror[:, 2:4] *= -1
    a = runner.track(ring, pool, flags=F64 | ufo.KICK, turns=50, parameters=C6, where=[-1])
    b = runner.track(ring, mirror, flags=F64 | ufo.KICK, turns=50, parameters=C6, where=[-1])
    assert np.array_equal(a[:, :, 0], b[:, :, 0]) and np.array_equal(a[:, :, 2], -b[:, :, 2])


def test_time_reversal_returns_to_start():
    """Track a sextupole FODO forward, flip the momenta, track the reversed
    line: the start point comes back to rounding level."""
    cell = lattices.fodo_cell()
    fwd = ufo.Line("F", [cell] * 8)
    rev = ufo.Line("R", list(reversed(fwd.flatten())))
    p = np.array([[1.5e-3, 1e-5, -7e-4, 2e-5, 0.0]])
    a = runner.track(fwd, p, flags=F64 | ufo.KICK | ufo.FIVED, turns=1, parameters=C6, where=[-1])[0, 0]
    q = np.array([[a[0], -a[1], a[2], -a[3], 0.0]])
    b = runner.track(rev, q, flags=F64 | ufo.KICK | ufo.FIVED, turns=1, parameters=C6, where=[-1])[0, 0]
    np.testing.assert_allclose([b[0], -b[1], b[2], -b[3]], p[0, :4], rtol=0, atol=1e-17)
