"""Synthetic lattices of the tracking workloads (FP64 teapot tracking).

``fodo_ring(cells)``
    cells x [QF L=0.5 m k1=+1.2 m^-2 (4 slices), SF thin sextupole
    k2L=+0.8 m^-2, drift 2.0 m, QD L=0.5 m k1=-1.2 m^-2 (4 slices), SD thin
    sextupole k2L=-1.4 m^-2, drift 2.0 m], then an aperture |x|,|y| < 40 mm
    checked once per turn.  The thin sextupoles are Multipoles: a zero-length
    Sextupole would kick with zero strength (the teapot scales by L/n).

``dba_ring(periods)``
    periods x a 16.9 m double-bend period of 25 elements: 2 combined-function
    sector bends (L=1.384 m, angle 2 pi/32, k1=-0.4 m^-2, 8 slices), 6
    quadrupoles (L 0.25-0.5 m, |k1| 1.0-2.2 m^-2, 8 slices), 4 sextupoles
    (L=0.2 m, k2 = +-38 m^-3, 4 slices) and 13 drifts; an aperture
    |x|,|y| < 30 mm at the end of the ring.  Per-period tunes (0.4266,
    0.3311): the whole x in [-20, 20] mm, y in [0, 8] mm scan survives 1024
    turns.  ``sext_scale`` multiplies the four sextupole strengths: at 4
    (k2 = +-152 m^-3) the dynamic aperture falls inside the scan and about
    40 % of its particles are lost within 1024 turns, most of them in the
    first 50 (the lossy scan).
"""
from __future__ import annotations

import math

from .elements import Aperture, Drift, Multipole, Quadrupole, Sbend, Sextupole
from .lattice import Line


def fodo_cell(tag=""):
    qf = Quadrupole(f"QF{tag}", length=0.5, k1=1.2, slices=4)
    qd = Quadrupole(f"QD{tag}", length=0.5, k1=-1.2, slices=4)
    sf = Multipole(f"SF{tag}", knl=[0.0, 0.0, 0.8])
    sd = Multipole(f"SD{tag}", knl=[0.0, 0.0, -1.4])
    d = Drift(f"D{tag}", length=2.0)
    return Line(f"CELL{tag}", [qf, sf, d, qd, sd, d])


def fodo_ring(cells=16, aperture=0.04, distinct=False):
    """FODO ring; ``distinct`` gives every cell its own element objects and
    labels (QF0, QF1, ...) so each magnet can be addressed separately."""
    if distinct:
        body = [fodo_cell(str(i)) for i in range(cells)]
    else:
        cell = fodo_cell()
        body = [cell] * cells
    ap = Aperture("AP", window=f"(fabs(x) < {aperture} && fabs(y) < {aperture})")
    return Line("RING", body + [ap])


# Working point: per-period tunes (0.4266, 0.3311), so Q = (1.707, 1.324) for
# 4 periods and (6.826, 5.297) for 16, at least 0.17 from every integer and
# half-integer in both rings (at QF2 = QF3 = 1.5 the 4-period vertical tune
# sat at 0.972, where 50 um misalignments leave some seeds without a
# closed orbit near the axis).
DBA_QUADS = {"QF1": (0.25, 2.2), "QD1": (0.30, -1.0), "QF2": (0.50, 1.3),
             "QF3": (0.50, 1.3), "QD2": (0.30, -1.0), "QF4": (0.25, 2.2)}
DBA_SEXTS = {"SD1": -38.0, "SF1": 38.0, "SF2": 38.0, "SD2": -38.0}


def dba_period(tag="", sext_scale=1.0):
    """One 16.9 m period (mirror-symmetric layout, distinct family labels)."""
    ang = 2.0 * math.pi / 32.0

    def q(name):
        L, k1 = DBA_QUADS[name]
        return Quadrupole(f"{name}{tag}", length=L, k1=k1, slices=8)

    def s(name):
        return Sextupole(f"{name}{tag}", length=0.2, k2=DBA_SEXTS[name] * sext_scale, slices=4)

    def b(name):
        return Sbend(f"{name}{tag}", length=1.384, angle=ang, k1=-0.4, slices=8)

    def d(name, L):
        return Drift(f"{name}{tag}", length=L)

    els = [d("DSS", 4.066), q("QF1"), d("D1", 0.25), q("QD1"), d("D2", 0.35), b("B1"),
           d("D3", 0.30), s("SD1"), d("D4", 0.20), q("QF2"), d("D5", 0.20), s("SF1"), d("DC", 0.50),
           s("SF2"), d("D5B", 0.20), q("QF3"), d("D4B", 0.20), s("SD2"), d("D3B", 0.30), b("B2"),
           d("D2B", 0.35), q("QD2"), d("D1B", 0.25), q("QF4"), d("DSSB", 4.066)]
    return Line(f"PERIOD{tag}", els)


def dba_ring(periods=16, aperture=0.03, distinct=False, sext_scale=1.0):
    if distinct:
        body = [dba_period(str(i), sext_scale) for i in range(periods)]
    else:
        p = dba_period("", sext_scale)
        body = [p] * periods
    ap = Aperture("AP", window=f"(fabs(x) < {aperture} && fabs(y) < {aperture})")
    return Line("RING", body + [ap])


def da_grid(nx, ny, xlim=0.02, ymax=0.008):
    """Initial (x, y) of an nx x ny dynamic-aperture grid, x fastest."""
    import numpy as np
    x = np.linspace(-xlim, xlim, nx)
    y = np.linspace(0.0, ymax, ny)
    xx, yy = np.meshgrid(x, y)
    return np.stack([xx.ravel(), yy.ravel()], axis=1)


def gaussian_bunch(n, seed=0, sx=1e-3, spx=1e-4, sy=1e-3, spy=1e-4, sdp=1e-3):
    """Gaussian bunch [n, 5] columns x, px, y, py, dp."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = np.empty((n, 5))
    for j, s in enumerate((sx, spx, sy, spy, sdp)):
        out[:, j] = rng.normal(0.0, s, n)
    return out


ERROR_ORDERS = (2, 3, 4, 5)          # b3..b6 / a3..a6 in 1-based (European) naming
REFERENCE_RADIUS = 0.01              # m


def dba_error_model(periods=4, seeds=256, seed0=1000, sigma_align=50e-6, sigma_rel=1e-4,
                    count_range=(4096, 32768)):
    """Error seeds of the ``periods``-period DBA ring.

    Every quadrupole and sextupole gets per seed: dx, dy ~ N(0, sigma_align),
    and normal/skew multipole errors dk_n, dk_ns (n = 2..5, i.e. b3..b6 and
    a3..a6) with relative strength b_n ~ N(0, sigma_rel) of the main field at
    the reference radius R = 10 mm: dk_n = b_n K_m R^(m-n) n!/m!, K_m the
    magnet's main strength (k1, m = 1; k2, m = 2) in the same per-length units.
    Seed s draws from numpy.random.default_rng(seed0 + s); its particle count
    is uniform in ``count_range``.

    Returns (ring, lattice_parameters, seed_values [seeds, P], counts [seeds]).
    """
    import numpy as np
    ring = dba_ring(periods, distinct=True)
    magnets = [e for e in ring.flatten() if isinstance(e, (Quadrupole, Sextupole))]
    names = ["dx", "dy"] + [f"dk{n}" for n in ERROR_ORDERS] + [f"dk{n}s" for n in ERROR_ORDERS]
    params = [(m.label, nm) for m in magnets for nm in names]
    for m in magnets:       # make the dkN / dkNs attributes addressable
        m.dkn = [0.0] * (max(ERROR_ORDERS) + 1)
        m.dks = [0.0] * (max(ERROR_ORDERS) + 1)
    values = np.zeros((seeds, len(params)))
    counts = np.zeros(seeds, dtype=np.int64)
    for s in range(seeds):
        rng = np.random.default_rng(seed0 + s)
        row = []
        for m in magnets:
            order, K = (1, m.k1) if isinstance(m, Quadrupole) else (2, m.k2)
            row += list(rng.normal(0.0, sigma_align, 2))
            b = rng.normal(0.0, sigma_rel, 2 * len(ERROR_ORDERS))
            for j, n in enumerate(list(ERROR_ORDERS) * 2):
                row.append(b[j] * K * REFERENCE_RADIUS ** (order - n) * math.factorial(n) / math.factorial(order))
        values[s] = row
        counts[s] = rng.integers(count_range[0], count_range[1] + 1)
    return ring, params, values, counts


def seed_grid(n, xlim=0.02, ymax=0.008):
    """First n points (x fastest) of the smallest ~5:1 grid over the DA
    rectangle holding n points: the same grid law for every particle count."""
    import numpy as np
    nx = int(math.ceil(math.sqrt(5.0 * n)))
    ny = int(math.ceil(n / nx))
    return da_grid(nx, ny, xlim, ymax)[:n]


def dba_family_model(seeds=64, seed0=1000, **kw):
    """The 4-period error model of ``dba_error_model`` with 8 sextupole
    families: the ring's 4 sextupole families per period, split by period
    parity (periods 0 and 2 form family set A, periods 1 and 3 set B).  The
    family strengths k2 are per-seed parameter columns (equal for every seed)
    bound by label; the per-magnet errors stay bound by flat element index.

    Returns (ring, lattice_parameters, seed_values, counts, family_parameters).
    """
    import numpy as np
    ring_e, prm_e, vals_e, counts = dba_error_model(periods=4, seeds=seeds, seed0=seed0, **kw)
    periods = [dba_period(f"{'AB'[r % 2]}") for r in range(4)]
    ring = Line("RING", periods + [Aperture("AP", window="(fabs(x) < 0.03 && fabs(y) < 0.03)")])
    flat_e = ring_e.flatten()
    flat = ring.flatten()
    for e in flat:
        if isinstance(e, (Quadrupole, Sextupole)):
            e.dkn = [0.0] * (max(ERROR_ORDERS) + 1)
            e.dks = [0.0] * (max(ERROR_ORDERS) + 1)
    index = {e.label: i for i, e in enumerate(flat_e)}
    prm = [(index[label], name) for label, name in prm_e]      # errors by flat index
    fam = [(f"{n}{t}", "k2") for t in "AB" for n in ("SD1", "SF1", "SF2", "SD2")]
    fam_vals = [DBA_SEXTS[n] for t in "AB" for n in ("SD1", "SF1", "SF2", "SD2")]
    values = np.concatenate([vals_e, np.tile(fam_vals, (seeds, 1))], axis=1)
    return ring, prm + fam, values, counts, fam
