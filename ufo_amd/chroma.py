"""Linear chromaticity from an Optics result (ufo/chroma.py:11-54).

Natural part: each element with a length and a gradient k1 (quadrupoles and
combined-function bends) contributes -k1/(4 pi) times the integral of beta_x
over its length (+k1/(4 pi) of beta_y), with beta transported through the
thick magnet from the Twiss values sampled at its entrance (cos/sin in the
focusing plane, cosh/sinh in the defocusing one).  Sextupoles (k2) add
k2 * integral(D beta) with beta and D transported as through a drift.

The sextupole integral follows the reference term by term, including the
sign of its last term: D' * gamma * L^4 / 4 enters with a minus sign
(ufo/chroma.py:44-45), where the drift-transported integral has a plus.
"""
import numpy as np

FOUR_PI = 4.0 * np.pi


def _sampled(optics, name):
    """Per observation point: (element attribute or 0) and the Twiss
    arrays, in the order of ``optics.where``."""
    els = [e for _, e in optics.where]
    n = min(len(els), len(optics.bx))
    vals = np.array([float(getattr(e, name, 0.0) or 0.0) for e in els[:n]])
    lens = np.array([float(getattr(e, "length", 0.0) or 0.0) for e in els[:n]])
    tw = {k: np.asarray(getattr(optics, k))[:n] for k in ("ax", "ay", "bx", "by", "dx", "dpx")}
    return vals, lens, tw


def _beta_integral(beta, alpha, gamma, K, L, focusing):
    """Integral over [0, L] of beta(s) through a thick quadrupole of
    |k1| = K^2, in its focusing (cos) or defocusing (cosh) plane."""
    u = K * L
    if focusing:
        half_sin2 = np.sin(2.0 * u) / (4.0 * K)
        sin2 = 1.0 - np.cos(u) ** 2
        return beta * (0.5 * L + half_sin2) - alpha * sin2 / K ** 2 + gamma * (0.5 * L - half_sin2) / K ** 2
    half_sinh2 = np.sinh(2.0 * u) / (4.0 * K)
    sinh2 = np.cosh(u) ** 2 - 1.0
    return beta * (0.5 * L + half_sinh2) - alpha * sinh2 / K ** 2 - gamma * (0.5 * L - half_sinh2) / K ** 2


def chromaticity(optics):
    """((natural dQx, natural dQy), (total dQx, total dQy)) of the elements
    listed in ``optics.where`` (Twiss values at their entrances)."""
    k1, L, tw = _sampled(optics, "k1")
    gx = (1.0 + tw["ax"] ** 2) / tw["bx"]
    gy = (1.0 + tw["ay"] ** 2) / tw["by"]
    nat_x = nat_y = 0.0
    q = (L != 0.0) & (k1 != 0.0)
    for foc in (True, False):
        m = q & ((k1 > 0.0) == foc)
        if not m.any():
            continue
        K = np.sqrt(np.abs(k1[m]))
        # k1 > 0 focuses x: x takes the cos form, y the cosh form
        ix = _beta_integral(tw["bx"][m], tw["ax"][m], gx[m], K, L[m], foc)
        iy = _beta_integral(tw["by"][m], tw["ay"][m], gy[m], K, L[m], not foc)
        nat_x -= float(np.sum(ix * k1[m]))
        nat_y += float(np.sum(iy * k1[m]))

    k2, L2, _ = _sampled(optics, "k2")
    s = (L2 != 0.0) & (k2 != 0.0)
    sx = sy = 0.0
    if s.any():
        l, k = L2[s], k2[s]
        d, dp = tw["dx"][s], tw["dpx"][s]

        def plane(b, a, g):
            return (d * (b + (g * l / 3.0 - a) * l) + dp * l * (0.5 * b - l * (2.0 * a / 3.0 + 0.25 * g * l))) * l * k

        sx = float(np.sum(plane(tw["bx"][s], tw["ax"][s], gx[s])))
        sy = -float(np.sum(plane(tw["by"][s], tw["ay"][s], gy[s])))
    nat_x /= FOUR_PI
    nat_y /= FOUR_PI
    return (nat_x, nat_y), (nat_x + sx / FOUR_PI, nat_y + sy / FOUR_PI)
