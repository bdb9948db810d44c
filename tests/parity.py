"""Parity helpers shared by the GPU tests.

The checker is the oracle's strict-IEEE build of the restated reference
kernels.  Rounding noise is amplified by the dynamics, so the bar is set per
particle from its measured sensitivity:

  sensitivity  max of three differences against the strict oracle, all of
               them the reference's own arithmetic: its fast-math build
               (stand-in for -cl-fast-relaxed-math -cl-mad-enable) and two
               strict runs with every input column moved by one ulp, up and
               down (all coordinates: in uncoupled linear motion a change of
               x never reaches the y plane).
  regular      same lost/alive flags in those runs and sensitivity <= 3e-13:
               rounding is not amplified beyond 3e-13 over the run.

Bar: on regular particles the HIP result is within 1e-12 (relative error
``|a - b| / max(|b|, rms_c)``, rms_c of coordinate c over the whole run) of
the strict oracle with identical
lost/alive flags; over all particles with stable flags, the median and 90th
percentile of its error are no larger than twice those of the reference's
own sensitivity (plus 1e-14), i.e. HIP is as close to the reference as the
reference is to itself; and on every one of those particles (the tail) its
error is at most 4 times that particle's own sensitivity plus 1e-14, so a
defect on a few particles cannot hide behind the aggregates.  Two rounding
perturbations of one chaotic orbit grow in directions that differ at random,
so the ratio of two such errors is heavy-tailed; with the sensitivity taken
as the largest of three runs the worst ratio over the nonlinear workloads is
below 4 (DBA ring at 20 mm, 64 turns: 3.8; FODO-16, 4096 particles, 1000
turns: 0.13).  Where every particle is chaotic (the reference's own spread
near 1e-6 after 30 turns through a strong octupole) a fourth independent
run exceeds four times the largest of three on about one particle in a
hundred, and the factor there is 8.
"""
import numpy as np

REL_TOL = 1e-12
REGULAR_TOL = 3e-13
STAT_FACTOR = 2.0
STAT_FLOOR = 1e-14
TAIL_FACTOR = 4.0


BOUND = 1.0    # metres / radians: beyond it a coordinate has left any aperture


def rel_err(a, b, axis_rms=0):
    """Per-entry relative error with the rms of b along ``axis_rms`` as the
    floor of the denominator.  The rms is taken over the finite entries
    within BOUND: a diverging particle (the Track loss test is NaN only, so
    coordinates can grow to 1e300 first) would otherwise raise the floor,
    or overflow it to inf, and hide every other particle's error."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    fin = np.isfinite(b)
    bb = np.where(fin, b, 0.0)
    inside = np.where(fin & (np.abs(b) <= BOUND), b, 0.0)
    rms = np.sqrt(np.mean(inside ** 2, axis=axis_rms, keepdims=True))
    rms = np.where(rms > 0, rms, 1.0)
    den = np.maximum(np.abs(bb), rms)
    err = np.abs(np.where(fin, a, 0.0) - bb) / den
    return np.where(fin, err, 0.0)


def particle_err(a, b):
    """Worst relative error per particle over samples and x, px, y, py; the
    floor of the denominator is the rms of each coordinate over all
    particles and samples (a per-sample rms would blow up where every
    trajectory crosses zero at once, as in uncoupled linear motion)."""
    return np.max(rel_err(a[:, :, :4], b[:, :, :4], axis_rms=(0, 1)), axis=(1, 2))


def nan_flags(t):
    return np.isnan(np.asarray(t, dtype=np.float64)[:, :, 0])


def reference_runs(track_fn, pool, fp32=False):
    """strict and fast runs of the oracle and the list of its two
    one-ulp-perturbed strict runs (every input moved up, then down);
    ``track_fn(pool, mode)`` runs it.  The ulp is that of the run's
    precision (an FP64 ulp vanishes when an FP32 run rounds its inputs)."""
    strict = track_fn(pool, "strict")
    fast = track_fn(pool, "fast")
    dt = np.float32 if fp32 else np.float64
    a = np.array(pool, dtype=dt)
    perturbed = [track_fn(np.nextafter(a, dt(np.inf)), "strict"), track_fn(np.nextafter(a, dt(-np.inf)), "strict")]
    return strict, fast, perturbed


CHAOS_TOL = 1e-9


def loss_turn_runs(sa_fn, pool, fp32=False, track_fn=None):
    """Loss turns of the oracle (``sa_fn(pool, mode)``) and the particles
    whose loss turn is regular: the strict run, the fast-math run and the
    strict run from every input moved by one ulp (in the run's precision)
    agree.  Over thousands of turns a particle at the chaotic edge of the
    aperture can survive in all three runs by chance while its trajectories
    have long parted; with ``track_fn(pool, mode)`` (the same particles
    tracked turn by turn) a particle is regular only if, in addition, its
    sensitivity over the run (fast-math and one-ulp-moved tracks against the
    strict one) stays below CHAOS_TOL."""
    strict = sa_fn(pool, "strict")
    fast = sa_fn(pool, "fast")
    dt = np.float32 if fp32 else np.float64
    moved = np.nextafter(np.array(pool, dtype=dt), dt(np.inf))
    perturbed = sa_fn(moved, "strict")
    regular = (strict == fast) & (strict == perturbed)
    if track_fn is not None:
        t_strict = track_fn(pool, "strict")
        sens = np.maximum(particle_err(track_fn(pool, "fast"), t_strict),
                          particle_err(track_fn(moved, "strict"), t_strict))
        regular &= sens <= CHAOS_TOL
    return strict, regular


def compare_tracks(gpu, strict, fast, perturbed):
    """``perturbed``: one perturbed oracle run or a list of them."""
    runs = [fast] + (list(perturbed) if isinstance(perturbed, (list, tuple)) else [perturbed])
    ns = nan_flags(strict)
    stable = np.ones(len(ns), dtype=bool)
    sens = np.zeros(len(ns))
    for run in runs:
        stable &= np.all(ns == nan_flags(run), axis=1)
        sens = np.maximum(sens, particle_err(run, strict))
    e_gpu = particle_err(gpu, strict)
    regular = stable & (sens <= REGULAR_TOL)
    flags_ok = np.array_equal(nan_flags(gpu)[stable], ns[stable])
    tail = e_gpu[stable] / (sens[stable] + STAT_FLOOR / TAIL_FACTOR)
    r = {
        "regular": regular,
        "flags_ok": flags_ok,
        "max_err": float(np.max(e_gpu[stable])) if np.any(stable) else 0.0,
        "tail_ratio": float(np.max(tail)) if tail.size else 0.0,
        "tail_over": int(np.sum(tail > TAIL_FACTOR)),
        "worst_regular": float(np.max(e_gpu[regular])) if np.any(regular) else 0.0,
        "median_gpu": float(np.median(e_gpu[stable])) if np.any(stable) else 0.0,
        "median_ref": float(np.median(sens[stable])) if np.any(stable) else 0.0,
        "p90_gpu": float(np.percentile(e_gpu[stable], 90)) if np.any(stable) else 0.0,
        "p90_ref": float(np.percentile(sens[stable], 90)) if np.any(stable) else 0.0,
    }
    return r


def summary(r):
    return (f"regular {int(r['regular'].sum())}/{len(r['regular'])}, worst on regular {r['worst_regular']:.2e}; "
            f"median err hip {r['median_gpu']:.2e} vs reference self {r['median_ref']:.2e}; "
            f"p90 hip {r['p90_gpu']:.2e} vs {r['p90_ref']:.2e}; max err hip {r['max_err']:.2e}, "
            f"worst err/sensitivity {r['tail_ratio']:.2f} ({r['tail_over']} over {TAIL_FACTOR:g})")


def assert_parity(r, min_regular_frac=0.25, tail_factor=TAIL_FACTOR):
    """The bar (module docstring).  ``tail_factor`` bounds each particle's
    error by that multiple of its own sensitivity (plus 1e-14)."""
    assert r["flags_ok"], "lost/alive flags differ on particles whose flags are stable"
    assert r["worst_regular"] <= REL_TOL, summary(r)
    assert r["tail_ratio"] <= tail_factor, summary(r)
    assert r["median_gpu"] <= STAT_FACTOR * r["median_ref"] + STAT_FLOOR, summary(r)
    assert r["p90_gpu"] <= STAT_FACTOR * r["p90_ref"] + STAT_FLOOR, summary(r)
    assert r["regular"].mean() >= min_regular_frac, summary(r)
