"""Synchrotron-radiation integrals and equilibrium emittance from an Optics
result (ufo/radiation.py:12-83).

For every bend in ``optics.where`` (angle != 0; length L, gradient k1, bend
radius rho = L/angle, Twiss and dispersion at its entrance) the integrals
over the magnet are evaluated in closed form for a combined-function sector
bend (the SLAC-1193 expressions, edge focusing ignored):

    I1 = sum <D> L / rho          I2 = sum L / rho^2
    I4 = sum <D> (1/rho^2 + 2 k1) L / rho
    I5 = sum <H> L / |rho|^3

with <D> and <H> the averages of the dispersion and of the curly-H function
over the magnet, D and the Twiss functions transported with the gradient k1
alone (the 1/rho^2 weak focusing is left out, as in the reference,
ufo/radiation.py:57-64).  Near k1 = 0 the factors are summed as series (the
reference evaluates 0/0 at k1 = 0 and returns NaN).
"""
import math

import numpy as np

from .constants import quantum_radiation_constant, radiation_loss_constant
from .lattice import Beam


class RadiationIntegrals:
    """I1, I2, I4, I5 with the damping partitions Jx = 1 - I4/I2 and
    Jt = 1 + I4/I2; ``U0`` (GeV per turn) and ``Ex`` (m rad) for the default
    Beam (3 GeV electrons), or for another beam through ``u0(beam)`` and
    ``ex(beam)``."""

    def __init__(self, I1, I2, I4, I5):
        self.I1 = I1
        self.I2 = I2
        self.I4 = I4
        self.I5 = I5
        self.Jx = 1.0 - I4 / I2
        self.Jt = 1.0 + I4 / I2

    def u0(self, beam=None):
        beam = beam if beam is not None else Beam()
        return radiation_loss_constant * (beam.energy * 1e-9) ** 4 * self.I2

    def ex(self, beam=None):
        beam = beam if beam is not None else Beam()
        return quantum_radiation_constant * beam.gamma ** 2 * self.I5 / self.I2 / self.Jx

    U0 = property(lambda self: self.u0())
    Ex = property(lambda self: self.ex())


def _factors(t):
    """The five functions of t = k1 L^2 the closed forms are built from,
    u = sqrt|t|, (C, S) = (cos u, sin u) for t > 0, (cosh u, sinh u) for
    t < 0:

        A0 = S/u              A1 = (1 - C)/t        A2 = (u - S)/(u t)
        B1 = (3u - 4S + SC)/(2u t^2)                B2 = (u - CS)/(2u t)

    For |t| < 0.5 they are summed as power series in -t (exact at t = 0,
    and free of the cancellation the closed forms suffer as k1 -> 0)."""
    t = np.asarray(t, dtype=np.float64)
    out = [np.empty_like(t) for _ in range(5)]
    small = np.abs(t) < 0.5
    if small.any():
        x = -t[small]
        fact = [math.factorial(i) for i in range(40)]
        terms = range(14)
        out[0][small] = sum(x ** m / fact[2 * m + 1] for m in terms)
        out[1][small] = sum(x ** m / fact[2 * m + 2] for m in terms)
        out[2][small] = sum(x ** m / fact[2 * m + 3] for m in terms)
        out[3][small] = 0.5 * sum(x ** m * (4.0 ** (m + 2) - 4.0) / fact[2 * m + 5] for m in terms)
        out[4][small] = 0.5 * sum(x ** m * 4.0 ** (m + 1) / fact[2 * m + 3] for m in terms)
    big = ~small
    if big.any():
        tb = t[big]
        u = np.sqrt(np.abs(tb))
        C = np.where(tb > 0.0, np.cos(u), np.cosh(u))
        S = np.where(tb > 0.0, np.sin(u), np.sinh(u))
        out[0][big] = S / u
        out[1][big] = (1.0 - C) / tb
        out[2][big] = (u - S) / (u * tb)
        out[3][big] = (3.0 * u - 4.0 * S + S * C) / (2.0 * u * tb * tb)
        out[4][big] = (u - C * S) / (2.0 * u * tb)
    return out


def _bend_terms(L, angle, k, d, dp, alpha, beta):
    """(<D> L, <H> L) of combined-function sector bends (one per array
    entry): D and the Twiss functions transported from the entrance values
    with horizontal focusing k1 (the 1/rho^2 term left out, as in the
    reference) and dispersion source 1/rho."""
    rho = L / angle
    gamma = (1.0 + alpha * alpha) / beta
    A0, A1, A2, B1, B2 = _factors(k * L * L)
    mean_d = d * A0 + dp * L * A1 + L * L / rho * A2
    h = gamma * d * d + 2.0 * alpha * d * dp + beta * dp * dp
    h = h + 2.0 * angle * ((alpha * d + beta * dp) * A1 - (gamma * d + alpha * dp) * L * A2)
    h = h + angle * angle * (gamma * L * L * B1 - alpha * L * A1 * A1 + beta * B2)
    return mean_d * L, h * L


def emittance(optics):
    """RadiationIntegrals of the bends listed in ``optics.where``."""
    els = [e for _, e in optics.where]
    n = min(len(els), len(optics.bx))

    def attr(name):
        return np.array([float(getattr(e, name, 0.0) or 0.0) for e in els[:n]])

    angle = attr("angle")
    m = angle != 0.0
    if not m.any():
        return RadiationIntegrals(0.0, 0.0, 0.0, 0.0)
    L, k = attr("length")[m], attr("k1")[m]
    th = angle[m]
    ax, bx = np.asarray(optics.ax)[:n][m], np.asarray(optics.bx)[:n][m]
    d, dp = np.asarray(optics.dx)[:n][m], np.asarray(optics.dpx)[:n][m]
    rho = L / th
    dl, hl = _bend_terms(L, th, k, d, dp, ax, bx)
    I1 = float(np.sum(dl / rho))
    I2 = float(np.sum(L / rho ** 2))
    I4 = float(np.sum(dl * (1.0 / rho ** 2 + 2.0 * k) / rho))
    I5 = float(np.sum(hl / np.abs(rho) ** 3))
    return RadiationIntegrals(I1, I2, I4, I5)
