"""The HIP kernels against the reference's own OpenCL kernels on the same
MI355X: the element statements of the reference's templates wrapped in the
kernel bodies of ufo/track.py and ufo/sa.py, built by the device's OpenCL
compiler with the reference's default options (-cl-fast-relaxed-math
-cl-mad-enable), launched as the reference launches them
(oracle/opencl.py).  This is the comparison the reference's own
users would make: same inputs, the reference's kernels on the GPU they run
on.  Kernel times of both are printed."""
import os

import numpy as np
import pytest

import ufo_amd as ufo
from ufo_amd import lattices
from oracle import runner

from conftest import OPTICS_DIR
from parity import loss_turn_runs, nan_flags, particle_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ocl():
    from oracle import opencl
    if not opencl.available():
        pytest.skip("no OpenCL GPU device")
    return opencl


def test_fodo16_track_full_size_against_reference_opencl(ocl):
    """The reference's CPU workload at full size (FODO-16, 4096 particles,
    6D, 1000 turns, where=[-1]) through the reference's Track kernel on the
    GPU: identical lost/alive flags and every sample of every particle within
    1e-12 (relative, rms floor per coordinate) of the reference's own OpenCL
    result, the bar the reference's regular orbits are held to, here met by
    all 4096."""
    ring = lattices.fodo_ring(16)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    prm = ["x", "px", "y", "py", "dp"]
    bunch = lattices.gaussian_bunch(4096, seed=0)
    tr = ufo.Track(ring, flags=flags, turns=1000, particles=len(bunch), parameters=prm, where=[-1])
    tr.parameters[:] = bunch
    tr.run()
    hip = np.array(tr.tracks)
    ref, ms = ocl.track(ring, bunch, flags, 1000, prm, [-1], threads=len(bunch), reps=3)
    strict = runner.track(ring, bunch, flags=flags, turns=1000, parameters=prm, where=[-1])
    e = particle_err(hip, ref)
    e_ref = particle_err(ref, strict)
    print(f"reference OpenCL kernel {np.median(ms):.3f} ms; HIP vs reference OpenCL: median {np.median(e):.2e}, "
          f"p90 {np.percentile(e, 90):.2e}, max {e.max():.2e}; reference OpenCL vs strict CPU oracle: median "
          f"{np.median(e_ref):.2e}, max {e_ref.max():.2e}")
    assert np.array_equal(nan_flags(hip), nan_flags(ref))
    assert e.max() <= 1e-12


def test_lossy_scan_loss_turns_against_reference_opencl(ocl):
    """The lossy DA scan's ring (sextupoles x4), 1024 points of the grid
    law, 1024 turns, through the reference's StableAperture kernel on the
    GPU (256 work-items claiming particles with atomic_inc): identical loss
    turns on every particle whose loss turn is regular (the CPU oracle's
    strict, fast-math and one-ulp-moved runs agree and their trajectories
    stay within 1e-9)."""
    turns = 1024
    ring = lattices.dba_ring(16, sext_scale=4.0)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    pool = lattices.da_grid(64, 16)
    sa = ufo.StableAperture(ring, flags=flags, turns=turns, particles=len(pool), parameters=["x", "y"])
    sa.parameters[:] = pool
    sa.run()
    ref, ms = ocl.stable_aperture(ring, pool, flags, turns, ["x", "y"], threads=256)
    strict, regular = loss_turn_runs(
        lambda p, m: runner.stable_aperture(ring, p, flags=flags, turns=turns, parameters=["x", "y"], mode=m),
        pool,
        track_fn=lambda p, m: runner.track(ring, p, flags=flags, turns=turns, parameters=["x", "y"], where=[-1],
                                           mode=m))
    hip = sa.lost.astype(np.int64)
    ref = ref.astype(np.int64)
    print(f"reference OpenCL kernel {ms[0]:.1f} ms; lost: HIP {np.sum(hip < turns)}, reference OpenCL "
          f"{np.sum(ref < turns)}; regular {regular.sum()}/{len(pool)}; HIP == reference OpenCL on "
          f"{np.sum(hip == ref)} of all, {np.sum(hip[regular] == ref[regular])} of the regular")
    assert regular.mean() > 0.7
    assert np.array_equal(hip[regular], ref[regular])


def test_da_scan_against_reference_opencl(ocl):
    """The headline DA scan (16-period DBA ring, 1024 x 1024 grid, 1024
    turns) through the reference's StableAperture kernel on the same GPU
    with 524 288 work-items (the best of 131 072 and 524 288): the same loss
    turns for every grid point (all survive); both kernel times printed."""
    import torch
    turns = 1024
    ring = lattices.dba_ring(16)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    pool = lattices.da_grid(1024, 1024)
    sa = ufo.StableAperture(ring, flags=flags, turns=turns, particles=len(pool), parameters=["x", "y"])
    sa.parameters[:] = pool
    sa.run()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    sa.launch()
    e1.record()
    torch.cuda.synchronize()
    hip_ms = e0.elapsed_time(e1)
    ref, ms = ocl.stable_aperture(ring, pool, flags, turns, ["x", "y"], threads=524288, reps=2)
    print(f"DA scan: HIP {hip_ms:.1f} ms, reference OpenCL {min(ms):.1f} ms ({min(ms) / hip_ms:.2f}x)")
    assert np.array_equal(ref, sa.lost)


def test_alba_thick_track_against_reference_opencl(ocl):
    """Thick ALBA (FIVED, FP64, dp = 1e-3), 64 particles, 50 turns, one
    observation point inside a thick quadrupole and one at the end: the
    reference's Track kernel on the GPU calls the OpenCL sin/cos/sinh/cosh
    of the platform (the HIP kernel folds them correctly rounded); identical
    flags, and the HIP result as close to the reference's OpenCL one as the
    reference's OpenCL result is to the strict CPU oracle (medians within a
    factor 2 + 1e-14)."""
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    quad = alba.RING.find(lambda e: isinstance(e, ufo.Quadrupole))[3]
    rng = np.random.default_rng(1)
    pool = np.stack([rng.normal(0, 2e-3, 64), rng.normal(0, 1e-4, 64),
                     rng.normal(0, 1e-3, 64), rng.normal(0, 1e-4, 64)], 1)
    flags = ufo.DOUBLE_PRECISION | ufo.FIVED
    prm = ["x", "px", "y", "py"]
    where = [quad + 0.5, -1]
    tr = ufo.Track(alba.RING, flags=flags, turns=50, particles=len(pool), parameters=prm, where=where, dp=1e-3)
    tr.parameters[:] = pool
    tr.run()
    hip = np.array(tr.tracks)
    ref, _ = ocl.track(alba.RING, pool, flags, 50, prm, where, dp=1e-3, threads=len(pool))
    strict = runner.track(alba.RING, pool, flags=flags, turns=50, parameters=prm, where=where, dp=1e-3)
    e, e_ref = particle_err(hip, ref), particle_err(ref, strict)
    print(f"ALBA thick: HIP vs reference OpenCL median {np.median(e):.2e} max {e.max():.2e}; reference OpenCL vs "
          f"strict CPU oracle median {np.median(e_ref):.2e} max {e_ref.max():.2e}")
    assert np.array_equal(nan_flags(hip), nan_flags(ref))
    assert np.median(e) <= 2 * np.median(e_ref) + 1e-14


def _against_reference_variants(ocl, hip, variants, turns):
    """Loss turns of HIP against the reference's default build on the
    particles whose loss turn is regular over all six reference builds, and
    the same comparison for each reference build against the default one on
    the particles regular over the other five (leave one out).  Returns
    (HIP mismatch rate, largest leave-one-out rate of the reference builds,
    report)."""
    ref = variants["default"][0]
    regular = ocl.regular_mask(variants, turns)
    mm = regular & (hip != ref)
    rate = mm.sum() / max(1, regular.sum())
    lines = [f"regular {regular.sum()} of {len(ref)}; HIP differs from the default build on {mm.sum()} "
             f"({rate:.2e}; later {np.sum(hip[mm] > ref[mm])}, earlier {np.sum(hip[mm] < ref[mm])}), "
             f"on {np.sum(hip != ref)} of all"]
    loo = {}
    for name in variants:
        if name == "default":
            continue
        reg = ocl.regular_mask(variants, turns, leave_out=(name,))
        m = reg & (variants[name][0] != ref)
        loo[name] = m.sum() / max(1, reg.sum())
        lines.append(f"  reference build '{name}' (left out): differs on {m.sum()} of {reg.sum()} ({loo[name]:.2e})")
    return rate, max(loo.values()), "\n".join(lines)


def test_lossy_scan_full_grid_against_reference_opencl(ocl):
    """The lossy DA scan at full size (sextupoles x4, 1024 x 1024 grid, 1024
    turns, 40 % lost) against the reference's StableAperture kernel on the
    GPU in six builds of its arithmetic (default and strict options, inputs
    moved one ulp up and down, strengths moved one ulp, whole-element drifts
    fused).  At the edge of the dynamic aperture a particle can escape by
    slow diffusion, and its loss turn then hangs on how each build rounds
    even where all six builds agree: each reference build, left out of the
    agreement, differs from the default build on ~2.5e-4 of the particles
    the other five agree on (the build with moved strengths 1.2e-3).  The
    bar: on the particles all six agree on, HIP differs from the default
    build no more often than twice the largest of those rates, i.e. HIP's
    arithmetic is one more build of the reference's."""
    turns = 1024
    ring = lattices.dba_ring(16, sext_scale=4.0)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    pool = lattices.da_grid(1024, 1024)
    sa = ufo.StableAperture(ring, flags=flags, turns=turns, particles=len(pool), parameters=["x", "y"])
    sa.parameters[:] = pool
    sa.run()
    variants = ocl.reference_variants(ring, pool, flags, turns, ["x", "y"], threads=524288)
    rate, ref_rate, report = _against_reference_variants(ocl, sa.lost.astype(np.int64), variants, turns)
    print(f"reference OpenCL {variants['default'][2]:.0f} ms; " + report)
    assert rate <= 2 * ref_rate


def test_error_seeds_against_reference_opencl(ocl):
    """The error-seed workload, 64 of its 256 seeds (1.2 M particles, 2048
    turns), against the reference's own way of tracking lattice variants:
    one StableAperture kernel with every seed's errors (400 per-particle
    parameter columns: dx, dy, b3-b6, a3-a6 of each quadrupole and
    sextupole) bound per particle (HOWTO.md:20-22), in the six builds of the
    lossy-scan test and with its bar."""
    from ufo_amd.batch import BatchedAperture
    turns = 2048
    ring, prm, vals, counts = lattices.dba_error_model(periods=4, seeds=256)
    seeds = 64
    vals, counts = vals[:seeds], counts[:seeds]
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    ba = BatchedAperture(ring, flags, turns, prm, vals, counts, block_turns=64)
    grid = np.concatenate([lattices.seed_grid(int(c)) for c in counts])
    ba.parameters[:] = grid
    ba.run()
    seed_of = np.repeat(np.arange(seeds), counts)
    pool = np.concatenate([grid, vals[seed_of]], axis=1)
    variants = ocl.reference_variants(ring, pool, flags, turns, ["x", "y"] + prm, threads=524288)
    rate, ref_rate, report = _against_reference_variants(ocl, ba.lost.astype(np.int64), variants, turns)
    print(f"{len(pool)} particles, {len(prm)} parameter columns, reference OpenCL {variants['default'][2]:.0f} ms; "
          + report)
    assert rate <= 2 * ref_rate


def test_dba_track_final_state_against_reference_opencl(ocl):
    """The headline ring's coordinates, not only its loss turns: 8192 points
    spanning the DA grid (128 x 64), 1024 turns, the HIP Track kernel's last
    sample against the final state of the reference's StableAperture kernel
    on the GPU in its six builds (``reference_variants``).  Errors relative
    to max(|coordinate|, rms of the coordinate over the survivors).  Over
    1024 turns of this strongly nonlinear ring one-ulp differences grow to
    ~1e-11 typically and ~1e-9 at the edge in the reference's own builds
    (the strict build against the default: median 9.3e-12); the bar: on the
    particles all six builds agree on within 1e-9, HIP's median distance to
    the default build is at most twice the largest median of the other
    builds, and on every survivor HIP is within 4 x the largest distance of
    any reference build from the default build (the per-particle
    sensitivity, ufo/sa.py:110-122) + 1e-15."""
    turns = 1024
    ring = lattices.dba_ring(16)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    pool = lattices.da_grid(128, 64)
    tr = ufo.Track(ring, flags=flags, turns=turns, particles=len(pool), parameters=["x", "y"], where=[-1])
    tr.parameters[:] = pool
    tr.run()
    hip = np.array(tr.tracks[:, -1, :4], dtype=np.float64)
    variants = ocl.reference_variants(ring, pool, flags, turns, ["x", "y"], threads=len(pool))
    ref_lost, ref, _ = variants["default"]
    alive = ref_lost == turns
    assert np.array_equal(np.isfinite(hip[:, 0]), alive)
    rms = np.sqrt(np.mean(ref[alive] ** 2, axis=0))
    scale = np.maximum(np.abs(ref), rms)

    def dist(state):
        return np.max(np.abs(state - ref) / scale, axis=1)

    regular = ocl.regular_mask(variants, turns)
    e_hip = dist(hip)
    spread = {name: dist(v[1]) for name, v in variants.items() if name != "default"}
    sens = np.max(np.stack(list(spread.values())), axis=0)
    worst_median = max(np.median(e[regular]) for e in spread.values())
    ratio = e_hip[alive] / np.maximum(sens[alive], 1e-300)
    print(f"regular {regular.sum()} of {alive.sum()} survivors; HIP vs default build: median "
          f"{np.median(e_hip[regular]):.2e}, p90 {np.percentile(e_hip[regular], 90):.2e}, max "
          f"{e_hip[regular].max():.2e}; within 1e-12: HIP {np.sum(e_hip[regular] <= 1e-12)}, "
          + ", ".join(f"{n} {np.sum(e[regular] <= 1e-12)} (median {np.median(e[regular]):.2e})"
                      for n, e in spread.items())
          + f"; HIP / sensitivity: median {np.median(ratio):.3f}, max {ratio.max():.2f}")
    assert regular.mean() > 0.9
    assert np.median(e_hip[regular]) <= 2 * worst_median
    assert np.all(e_hip[alive] <= 4 * sens[alive] + 1e-15)


@pytest.mark.parametrize("case", ["fodo16_kick", "fodo16_thick", "dba_kick_fived"])
def test_fp32_track_against_reference_opencl(ocl, case):
    """The reference's default precision (no DOUBLE_PRECISION flag: FP32
    state and constants, ``-cl-single-precision-constant``): Track through
    the reference's OpenCL kernel on the GPU.  In FP32 the rounding of every
    operation moves trajectories by ~1e-6 per turn, so over these runs two
    correct builds part by ~1e-4-1e-3: the bar is identical lost/alive flags
    and a median distance between HIP and the reference kernel at most twice
    the larger median distance of two of the reference's own builds from the
    strict CPU oracle (its OpenCL kernel, and the oracle built with fast-math
    contraction)."""
    ring, flags, turns, prm, pool = {
        "fodo16_kick": (lattices.fodo_ring(16), ufo.KICK, 1000, ["x", "px", "y", "py", "dp"],
                        lattices.gaussian_bunch(4096, seed=0)),
        "fodo16_thick": (lattices.fodo_ring(16), 0, 1000, ["x", "px", "y", "py", "dp"],
                         lattices.gaussian_bunch(4096, seed=0)),
        "dba_kick_fived": (lattices.dba_ring(16), ufo.KICK | ufo.FIVED, 200, ["x", "y"], lattices.da_grid(64, 64)),
    }[case]
    tr = ufo.Track(ring, flags=flags, turns=turns, particles=len(pool), parameters=prm, where=[-1])
    tr.parameters[:] = pool
    tr.run()
    hip = np.array(tr.tracks, dtype=np.float64)
    ref, _ = ocl.track(ring, pool, flags, turns, prm, [-1], threads=len(pool))
    ref = ref.astype(np.float64)
    strict, fast = (runner.track(ring, pool, flags=flags, turns=turns, parameters=prm, where=[-1],
                                 mode=m).astype(np.float64) for m in ("strict", "fast"))
    e = np.nanmedian(particle_err(hip, ref))
    spread = {"reference OpenCL": np.nanmedian(particle_err(ref, strict)),
              "fast-math oracle": np.nanmedian(particle_err(fast, strict))}
    print(f"{case}: HIP vs reference OpenCL median {e:.2e}; from the strict oracle: "
          + ", ".join(f"{k} {v:.2e}" for k, v in spread.items()))
    assert np.array_equal(nan_flags(hip), nan_flags(ref))
    assert e <= 2 * max(spread.values())
