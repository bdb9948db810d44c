"""Chromaticity, radiation integrals and resonance driving terms
(ufo/chroma.py, ufo/radiation.py, ufo/rdt.py) computed from the oracle's
Optics, pinned by direct numerical integration: the closed forms against an
ODE integration of the same model, the Twiss-weighted integrals against
Optics sampled inside the magnets (fractional ``where``), and the driving
terms against the period structure of a 16-period ring."""
import json
import os

import numpy as np
import pytest
from scipy.integrate import simpson, solve_ivp

import ufo_amd as ufo
from ufo_amd import lattices, radiation
from oracle import runner

from conftest import OPTICS_DIR, ROOT

GOLDEN_VALUES = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_values.json")))

FL = ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION


def _alba():
    return ufo.Lattice(path=os.path.join(OPTICS_DIR, "alba.mad"))


def _integrate(y, x):
    """Simpson's rule on n = 2^k + 1 equally spaced samples, Richardson
    extrapolated with the rule on every second sample (error O(h^6))."""
    fine = simpson(y, x=x)
    coarse = simpson(y[::2], x=x[::2])
    return fine + (fine - coarse) / 15.0


def _sampled_optics(line, idxs, n=17):
    """Optics at n equally spaced points through each flat element of idxs,
    in one run: returns f [n] and, per element, a dict of the Twiss arrays
    at those points."""
    f = np.linspace(0.0, 1.0, n)
    where = []
    for idx in idxs:
        where += [float(idx + x) for x in f[:-1]] + [idx + 1]
    o = runner.Optics(line, where=where, flags=FL)
    o.run()
    out = []
    for j in range(len(idxs)):
        sl = slice(j * n, (j + 1) * n)
        out.append({k: np.asarray(getattr(o, k))[sl] for k in ("ax", "ay", "bx", "by", "dx", "dpx")})
    return f, out


@pytest.mark.parametrize("k", [0.8, -0.57, 0.0, 1e-9, -2.0])
def test_bend_closed_forms_match_ode(k):
    """<D> L and <H> L of a sector bend in the model of the closed forms
    (horizontal focusing k1, dispersion source 1/rho), for random entrance
    Twiss/dispersion, against a direct integration of
    D'' = -k1 D + 1/rho and M' = [[0, 1], [-k1, 0]] M."""
    rng = np.random.default_rng(int(abs(k) * 1000) + 3)
    L, angle = 1.384, 2 * np.pi / 32
    rho = L / angle
    for _ in range(3):
        a, b = rng.normal(0, 1), rng.uniform(0.5, 8.0)
        g = (1 + a * a) / b
        d, dp = rng.normal(0, 0.2), rng.normal(0, 0.05)

        def rhs(s, y):
            D, Dp, m11, m12, m21, m22 = y
            return [Dp, -k * D + 1.0 / rho, m21, m22, -k * m11, -k * m12]

        s = np.linspace(0.0, L, 4001)
        sol = solve_ivp(rhs, (0.0, L), [d, dp, 1.0, 0.0, 0.0, 1.0], t_eval=s, rtol=1e-12, atol=1e-14,
                        method="DOP853")
        D, Dp, m11, m12, m21, m22 = sol.y
        beta = m11 ** 2 * b - 2 * m11 * m12 * a + m12 ** 2 * g
        alpha = -m11 * m21 * b + (m11 * m22 + m12 * m21) * a - m12 * m22 * g
        gamma = m21 ** 2 * b - 2 * m21 * m22 * a + m22 ** 2 * g
        H = gamma * D ** 2 + 2 * alpha * D * Dp + beta * Dp ** 2
        dl, hl = radiation._bend_terms(np.array([L]), np.array([angle]), np.array([k]), np.array([d]),
                                       np.array([dp]), np.array([a]), np.array([b]))
        assert dl[0] == pytest.approx(simpson(D, x=s), rel=1e-9, abs=1e-12)
        assert hl[0] == pytest.approx(simpson(H, x=s), rel=1e-9, abs=1e-14)


def test_radiation_integrals_alba():
    """ALBA (tests/radiation.py): I2 is geometry and equals the MAD-X value;
    the dispersion integral sampled through every bend reproduces MAD-X's
    momentum compaction; the closed forms, which transport D and beta with
    k1 alone (no 1/rho^2 term, ufo/radiation.py:57-64), stay within 2% of
    the sampled integrals."""
    g = GOLDEN_VALUES["alba_radiation_madx"]
    alba = _alba()
    ring = alba.RING
    flat = ring.flatten()
    bends = ring.find(lambda e: hasattr(e, "angle"))
    o = runner.Optics(ring, where=bends, flags=FL)
    o.run()
    r = ufo.emittance(o)
    assert r.I2 == pytest.approx(g["I2"], rel=1e-9)
    f, samples = _sampled_optics(ring, bends, 9)
    i1 = i5 = 0.0
    for i, sm in zip(bends, samples):
        e = flat[i]
        rho = e.length / e.angle
        x = f * e.length
        gx = (1 + sm["ax"] ** 2) / sm["bx"]
        H = gx * sm["dx"] ** 2 + 2 * sm["ax"] * sm["dx"] * sm["dpx"] + sm["bx"] * sm["dpx"] ** 2
        i1 += _integrate(sm["dx"], x) / rho
        i5 += _integrate(H, x) / abs(rho) ** 3
    assert i1 / ring.length == pytest.approx(g["alpha_c"], rel=1e-4)
    assert r.I1 == pytest.approx(i1, rel=0.02) and r.I5 == pytest.approx(i5, rel=0.02)
    assert r.Jx == pytest.approx(1 - r.I4 / r.I2) and r.Ex > 0
    assert r.U0 == pytest.approx(ufo.radiation_loss_constant * 3.0 ** 4 * r.I2)


def test_natural_chromaticity_of_quadrupoles_vs_sampled_beta():
    """The quadrupole closed form equals -k1/(4 pi) * integral of beta_x
    (+k1/(4 pi) beta_y) with beta sampled through each thick quadrupole."""
    alba = _alba()
    ring = alba.RING
    flat = ring.flatten()
    quads = ring.find(lambda e: isinstance(e, ufo.Quadrupole))
    o = runner.Optics(ring, where=quads, flags=FL)
    o.run()
    (nx, ny), _ = ufo.chromaticity(o)
    sx = sy = 0.0
    f, samples = _sampled_optics(ring, quads, 9)
    for i, sm in zip(quads, samples):
        e = flat[i]
        x = f * e.length
        sx -= e.k1 * _integrate(sm["bx"], x)
        sy += e.k1 * _integrate(sm["by"], x)
    assert nx == pytest.approx(sx / (4 * np.pi), rel=1e-7)
    assert ny == pytest.approx(sy / (4 * np.pi), rel=1e-7)
    assert nx < 0


def test_sextupole_chromaticity_vs_sampled_dispersion():
    """The sextupole term equals k2/(4 pi) * integral of D beta through each
    sextupole except for the reference's sign of the D' gamma L^4 / 4 term
    (ufo/chroma.py:44-45), which makes it smaller by k2 D' gamma L^4 / 2."""
    alba = _alba()
    ring = alba.RING
    flat = ring.flatten()
    sext = ring.find(lambda e: isinstance(e, ufo.Sextupole))
    o = runner.Optics(ring, where=sext, flags=FL)
    o.run()
    (nx, ny), (tx, ty) = ufo.chromaticity(o)
    assert nx == 0.0 and ny == 0.0
    num_x = num_y = quirk_x = quirk_y = 0.0
    f, samples = _sampled_optics(ring, sext, 5)
    for j, (i, sm) in enumerate(zip(sext, samples)):
        e = flat[i]
        x = f * e.length
        num_x += e.k2 * _integrate(sm["dx"] * sm["bx"], x)
        num_y -= e.k2 * _integrate(sm["dx"] * sm["by"], x)
        gx = (1 + o.ax[j] ** 2) / o.bx[j]
        gy = (1 + o.ay[j] ** 2) / o.by[j]
        quirk_x += e.k2 * o.dpx[j] * gx * e.length ** 4 / 2
        quirk_y -= e.k2 * o.dpx[j] * gy * e.length ** 4 / 2
    assert tx == pytest.approx((num_x - quirk_x) / (4 * np.pi), rel=1e-9, abs=1e-12)
    assert ty == pytest.approx((num_y - quirk_y) / (4 * np.pi), rel=1e-9, abs=1e-12)


def test_rdt_of_a_ring_is_the_period_sum_times_the_phase_series():
    """16 identical DBA periods: each driving term of the ring equals that
    of one period times sum_p exp(i n p mu_period), n the term's phase
    multiplier (3 mu_x, -mu_x, mu_x + 2 mu_y, -mu_x + 2 mu_y, -mu_x)."""
    period = ufo.Line("P", [lattices.dba_period()])
    ring = ufo.Line("R", [period] * 16)     # no aperture: the optics rays have unit amplitude
    o_r = runner.Optics(ring, where=ring.find(lambda e: isinstance(e, ufo.Sextupole)), flags=FL)
    o_p = runner.Optics(period, where=period.find(lambda e: isinstance(e, ufo.Sextupole)), flags=FL)
    o_r.run()
    o_p.run()
    mx, my = 2 * np.pi * o_p.qx, 2 * np.pi * o_p.qy
    fr, fp = ufo.rdt(o_r), ufo.rdt(o_p)
    p = np.arange(16)
    for f_ring, f_per, phase in zip(fr, fp, (3 * mx, -mx, mx + 2 * my, -mx + 2 * my, -mx)):
        expect = f_per * np.sum(np.exp(1j * p * phase))
        assert abs(f_ring - expect) <= 1e-9 * abs(f_per) * 16
    assert abs(fp[0]) > 1.0
