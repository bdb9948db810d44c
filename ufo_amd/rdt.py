"""First-order sextupole resonance driving terms from an Optics result
(ufo/rdt.py:11-28).

For every sextupole in ``optics.where`` (length L, k2, Twiss and phase
advance at its entrance), with w = k2 L sqrt(beta_x):

    f3000 = sum w beta_x exp(3i mu_x)       f1200 = sum w beta_x exp(-i mu_x)
    f1020 = sum w beta_y exp(i(mu_x + 2 mu_y))
    f0120 = sum w beta_y exp(i(-mu_x + 2 mu_y))
    f0111 = sum w beta_y exp(-i mu_x)

The sums carry no normalisation factors, as in the reference.
"""
import numpy as np


def rdt(optics):
    els = [e for _, e in optics.where]
    n = min(len(els), len(optics.bx))
    k2 = np.array([float(getattr(e, "k2", 0.0) or 0.0) for e in els[:n]])
    L = np.array([float(getattr(e, "length", 0.0) or 0.0) for e in els[:n]])
    m = (L != 0.0) & (k2 != 0.0)
    bx, by = np.asarray(optics.bx)[:n][m], np.asarray(optics.by)[:n][m]
    mx, my = np.asarray(optics.mux)[:n][m], np.asarray(optics.muy)[:n][m]
    w = k2[m] * L[m] * np.sqrt(bx)
    f3000 = np.sum(w * bx * np.exp(3j * mx))
    f1200 = np.sum(w * bx * np.exp(-1j * mx))
    f1020 = np.sum(w * by * np.exp(1j * (mx + 2.0 * my)))
    f0120 = np.sum(w * by * np.exp(1j * (2.0 * my - mx)))
    f0111 = np.sum(w * by * np.exp(-1j * mx))
    return tuple(complex(v) if m.any() else 0.0 for v in (f3000, f1200, f1020, f0120, f0111))
