"""Parity helpers shared by the GPU tests.

Bar: relative coordinate error ``|a - b| / max(|b|, rms_c)`` at most 1e-12
on regular surviving particles, identical lost/alive flags.  A particle is
regular when the oracle's strict-IEEE and fast-math builds (the reference's
own noise floor: its kernels are compiled with -cl-fast-relaxed-math
-cl-mad-enable) agree to 1e-13 on it.
"""
import numpy as np

REL_TOL = 1e-12
REGULAR_TOL = 1e-13


def rel_err(a, b, axis_rms=0):
    """Per-entry relative error with the rms of b along ``axis_rms`` as the
    floor of the denominator (finite entries only)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    fin = np.isfinite(b)
    bb = np.where(fin, b, 0.0)
    rms = np.sqrt(np.mean(bb ** 2, axis=axis_rms, keepdims=True))
    rms = np.where(rms > 0, rms, 1.0)
    den = np.maximum(np.abs(bb), rms)
    err = np.abs(np.where(fin, a, 0.0) - bb) / den
    return np.where(fin, err, 0.0)


def track_regular(strict, fast, tol=REGULAR_TOL):
    """Particles of a Track result [particles, samples, 6] whose strict and
    fast oracle runs agree to ``tol`` at every sample."""
    err = rel_err(strict[:, :, :4], fast[:, :, :4])
    same_nan = np.all(np.isnan(strict[:, :, 0]) == np.isnan(fast[:, :, 0]), axis=1)
    return same_nan & (np.max(err, axis=(1, 2)) <= tol)


def compare_tracks(gpu, strict, fast):
    """Returns (max rel err on regular particles, regular mask, nan flags
    equal on regular particles)."""
    reg = track_regular(strict, fast)
    err = rel_err(gpu[:, :, :4], strict[:, :, :4])
    flags_ok = np.array_equal(np.isnan(gpu[reg, :, 0]), np.isnan(strict[reg, :, 0]))
    worst = float(np.max(err[reg])) if np.any(reg) else 0.0
    return worst, reg, flags_ok
