"""GPU parity: HIP kernels against the oracle on the same seeded inputs.

Every test calls the product (ufo_amd.Track / StableAperture / Optics), which
generates HIP source, compiles it with hipRTC through libufo_hip.so and
launches it on cuda:0; the oracle (gcc build of the restated reference
kernels) is the checker.
"""
import os

import numpy as np
import torch
import pytest

import ufo_amd as ufo
from ufo_amd import lattices
from oracle import runner

from conftest import OPTICS_DIR
import host_emulation
from parity import assert_parity, compare_tracks, loss_turn_runs, reference_runs, summary

pytestmark = pytest.mark.gpu

COORD6 = ["x", "px", "y", "py", "dp"]


def _track_both(line, pool, flags, turns, parameters, where, dp=0.0):
    """HIP Track result and the three oracle runs (strict, fast-math, strict
    from every input moved by one ulp).  On teapot lattices the HIP result must also
    equal the host build of the emitted kernel bit for bit."""
    tr = ufo.Track(line, flags=flags, turns=turns, particles=len(pool), parameters=parameters, where=where, dp=dp)
    tr.parameters[:] = pool
    tr.run()
    gpu = np.array(tr.tracks)

    def oracle(p, mode):
        return runner.track(line, p, flags=flags, turns=turns, parameters=parameters, where=where, dp=dp,
                            mode=mode)

    refs = reference_runs(oracle, pool, fp32=not (flags & ufo.DOUBLE_PRECISION))
    if not _has_transcendentals(tr.kernel.src):
        host = host_emulation.run_track(tr.kernel, tr.parameters, tr.samples)
        same = (gpu == host) | (np.isnan(gpu) & np.isnan(host))
        assert same.all(), f"HIP differs from the host build of the emitted kernel in {np.sum(~same)} entries"
    return (gpu,) + refs


def _has_transcendentals(src):
    """Whether the kernel calls a device-library function whose rounding the
    host build does not reproduce.  The correctly rounded double-double
    functions of once-per-particle invariants (ufo_cr_*) are plain
    arithmetic and reproduce bit for bit, as do FP64 division and sqrt
    (correctly rounded on the device); FP32 division and sqrt are left out."""
    import re
    from ufo_amd import crmath
    body = src.replace(crmath.source(), "")
    calls = set(re.findall(r"\b([a-z_0-9]+)\(", body)) - {"if", "for", "unroll", "__launch_bounds__", "__builtin_nan",
                                                           "__builtin_nanf", "fma", "fabs", "isnan", "sqrt"}
    calls = {c for c in calls if not c.startswith(("ufo_cr_", "ufo_track", "ufo_sa", "ufo_optics"))}
    fp32 = "typedef float ufloat" in src
    return bool(calls) or (fp32 and (" / " in body or "sqrt(" in body))


def test_fodo_track_6d():
    """FODO-16 (4096-particle workload at reduced size), 6D, dp per particle."""
    ring = lattices.fodo_ring(16)
    bunch = lattices.gaussian_bunch(512, seed=0)
    bunch[:8, 0] = np.linspace(0.03, 0.06, 8)      # some lost at the aperture
    gpu, strict, fast, moved = _track_both(ring, bunch, ufo.KICK | ufo.DOUBLE_PRECISION, 100, COORD6, [-1])
    r = compare_tracks(gpu, strict, fast, moved)
    print(summary(r))
    assert_parity(r, 0.3)
    assert np.array_equal(np.isnan(gpu[:, :, 0]), np.isnan(strict[:, :, 0]))
    # dp and z columns are carried unchanged
    alive = ~np.isnan(gpu[:, -1, 0])
    assert np.all(gpu[alive, :, 5] == bunch[alive, None, 4])
    assert np.all(gpu[alive, :, 4] == 0.0)


def test_track_launch_overwrites_a_stale_samples_buffer():
    """Track.launch() on a samples buffer holding stale values (7.0): the
    kernel writes every sample of its turns, NaN after a particle's loss,
    and the driver fills the never-written tail, so the result equals run()
    bit for bit (particles lost part-way, a fractional point on a thin
    element)."""
    ring = lattices.fodo_ring(16)
    bunch = lattices.gaussian_bunch(300, seed=3)
    bunch[:6, 0] = [0.035, 0.038, 0.045, 0.05, 0.0399, -0.039]
    sf = ring.find(lambda e: isinstance(e, ufo.Multipole))[0]
    tr = ufo.Track(ring, flags=ufo.KICK | ufo.DOUBLE_PRECISION, turns=200, particles=len(bunch),
                   parameters=COORD6, where=[2.5, sf + 0.5, -1])
    tr.parameters[:] = bunch
    tr.run()
    want = np.array(tr.tracks)
    lost = np.isnan(want[:, 399, 0])
    assert 0 < lost.sum() < len(bunch) and np.all(np.isnan(want[:, 400:, :]))
    tr._tracks_dev.fill_(7.0)
    tr.launch()
    got = tr.tracks_device.cpu().numpy()
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("double", [True, False])
def test_track_run_split_over_launches(double):
    """Track.run splits the turns over launches (copying finished samples
    back meanwhile); one launch, 7 uneven pieces and one launch per turn give
    the same tracks bit for bit, with particles lost part-way and a
    fractional observation point on a thin element that is never written."""
    ring = lattices.fodo_ring(16)
    bunch = lattices.gaussian_bunch(300, seed=3)
    bunch[:6, 0] = [0.035, 0.038, 0.045, 0.05, 0.0399, -0.039]
    sf = ring.find(lambda e: isinstance(e, ufo.Multipole))[0]
    flags = ufo.KICK | (ufo.DOUBLE_PRECISION if double else 0)
    tr = ufo.Track(ring, flags=flags, turns=200, particles=len(bunch), parameters=COORD6,
                   where=[2.5, sf + 0.5, -1])
    tr.parameters[:] = bunch
    runs = []
    for chunks in (1, 7, 200):
        tr.copy_chunks = chunks
        tr.tracks[:] = 0.0
        tr.run()
        runs.append(np.array(tr.tracks))
    lost = np.isnan(runs[0][:, 399, 0])
    assert 0 < lost.sum() < len(bunch) and np.all(np.isnan(runs[0][:, 400:, :]))
    u = np.uint64 if double else np.uint32
    for r in runs[1:]:
        np.testing.assert_array_equal(r.view(u), runs[0].view(u))


def test_fodo16_4096_particles_1000_turns_full_size():
    """The reference's own CPU workload at full size: FODO-16, 4096
    particles (seed 0), 6D, 1000 turns, where=[-1].  After 1000 turns the
    reference's arithmetic disagrees with itself at the 4e-12 level (median
    over particles of its fast-math and one-ulp-moved runs), so few
    particles are regular: the bar is identical loss flags, 1e-12 on the
    regular ones, and the statistical comparison on all."""
    ring = lattices.fodo_ring(16)
    bunch = lattices.gaussian_bunch(4096, seed=0)
    gpu, strict, fast, moved = _track_both(ring, bunch, ufo.KICK | ufo.DOUBLE_PRECISION, 1000, COORD6, [-1])
    r = compare_tracks(gpu, strict, fast, moved)
    print(summary(r))
    assert_parity(r, min_regular_frac=0.0)
    assert np.array_equal(np.isnan(gpu[:, :, 0]), np.isnan(strict[:, :, 0]))


def test_dba_stable_aperture_subset():
    """DA ring, a 32x8 subset of the grid (including the lossy corner),
    1024 turns: identical loss turns on regular particles."""
    ring = lattices.dba_ring(16)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    pool = lattices.da_grid(32, 8)
    sa = ufo.StableAperture(ring, flags=flags, turns=1024, particles=len(pool), parameters=["x", "y"])
    sa.parameters[:] = pool
    sa.run()
    strict, regular = loss_turn_runs(
        lambda p, m: runner.stable_aperture(ring, p, flags=flags, turns=1024, parameters=["x", "y"], mode=m), pool)
    print(f"lost: {np.sum(strict < 1024)}; regular {regular.sum()}/{len(regular)}")
    assert np.array_equal(sa.lost[regular], strict[regular])


def test_lossy_dba_stable_aperture_subset():
    """The lossy DA scan (sextupoles x4, k2 = +-152): 1024 points of the
    scan's grid law (64 x 16 over x in [-20, 20] mm, y in [0, 8] mm), 1024
    turns, launched on 256 lanes (one workgroup) so that every lane refills
    from the work counter several times: identical loss turns on every particle whose loss
    turn is regular (the strict, fast-math and one-ulp-moved oracle runs
    agree and their trajectories stay within 1e-9 of each other), with loss
    turns spread over the run (ufo/sa.py:116-122)."""
    import torch
    turns = 1024
    ring = lattices.dba_ring(16, sext_scale=4.0)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    pool = lattices.da_grid(64, 16)
    sa = ufo.StableAperture(ring, flags=flags, turns=turns, particles=len(pool), parameters=["x", "y"])
    sa.parameters[:] = pool
    sa.pool_device.copy_(sa._pool_host)
    sa.launch(lanes=256)
    gpu = sa.lost_device.cpu().numpy().view(np.uint32).astype(np.int64)
    assert sa.lanes == 256
    strict, regular = loss_turn_runs(
        lambda p, m: runner.stable_aperture(ring, p, flags=flags, turns=turns, parameters=["x", "y"], mode=m),
        pool,
        track_fn=lambda p, m: runner.track(ring, p, flags=flags, turns=turns, parameters=["x", "y"], where=[-1],
                                           mode=m))
    lost = strict < turns
    print(f"lost {lost.sum()}/{len(pool)}, loss turns 10/50/90%: {np.percentile(strict[lost], [10, 50, 90])}, "
          f"max {strict[lost].max()}; regular {regular.sum()}; differ on regular "
          f"{np.sum(gpu[regular] != strict[regular])}, on all {np.sum(gpu != strict)}")
    assert 0.2 < lost.mean() < 0.7 and strict[lost].max() > 100
    assert regular.mean() > 0.7
    assert np.array_equal(gpu[regular], strict[regular])


def test_dba_track_coordinates():
    """DA ring tracked 64 turns from grid points.  The grid reaches 20 mm,
    where the motion is strongly nonlinear and rounding differences grow
    quickly: after 64 turns 15 of the 64 points are regular (the oracle's
    strict, fast-math and two one-ulp-moved runs agree to 3e-13)."""
    ring = lattices.dba_ring(16)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    pool = lattices.da_grid(16, 4)
    gpu, strict, fast, moved = _track_both(ring, pool, flags, 64, ["x", "y"], [-1])
    r = compare_tracks(gpu, strict, fast, moved)
    print(summary(r))
    assert_parity(r, min_regular_frac=0.2)


def test_alba_thick_track_fp64():
    """ALBA ring with thick quads, sector bends with pole faces, sextupoles:
    50 turns, FIVED, dp = 1e-3 literal."""
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    rng = np.random.default_rng(1)
    pool = np.stack([rng.normal(0, 2e-3, 64), rng.normal(0, 1e-4, 64),
                     rng.normal(0, 1e-3, 64), rng.normal(0, 1e-4, 64)], 1)
    flags = ufo.DOUBLE_PRECISION | ufo.FIVED
    gpu, strict, fast, moved = _track_both(alba.RING, pool, flags, 50, ["x", "px", "y", "py"], [-1], dp=1e-3)
    r = compare_tracks(gpu, strict, fast, moved)
    print(summary(r))
    assert_parity(r)


def _all_elements_line():
    q = ufo.Quadrupole("Q", length=0.6, k1=1.1, slices=5, dx=2e-4, dy=-1e-4, dkn=[0, 0, 3.0], dks=[0, 0.01, 0, 8.0])
    b = ufo.Sbend("B", length=1.2, angle=0.15, k1=-0.3, e1=0.05, e2=0.07, hgap=0.02, fint=0.5, slices=6)
    r = ufo.Rbend("R", length=0.9, angle=0.1, k1=0.0, slices=3)
    s = ufo.Sextupole("S", length=0.3, k2=20.0, k2s=5.0, dx=1e-4, dkn=[0, 0, 0, 4.0])
    o = ufo.Octupole("O", length=0.2, k3=300.0, slices=2)
    m = ufo.Multipole("M", knl=[1e-5, 0.02, 1.5, 40.0, 800.0, 3e4], ksl=[2e-5, 0.01, 0.5, 10.0, 0, 1e4],
                      dx=1e-4, dy=2e-4)
    w = ufo.Wire("W", x=0.05, y=0.03, k=1e-6)
    d1, d2 = ufo.Drift("D1", length=0.7), ufo.Drift("D2", length=1.3)
    qd = ufo.Quadrupole("QD", length=0.5, k1=-1.3, slices=4)
    ap = ufo.Aperture("AP", window="(fabs(x) < 0.02 && fabs(y) < 0.02)")
    mk = ufo.Marker("MK")
    cell = ufo.Line("C", [q, d1, b, d2, s, mk, qd, d1, o, m, r, d2, w, ap])
    return ufo.Line("RING", [cell] * 3)


@pytest.mark.parametrize("flags", [
    ufo.KICK | ufo.DOUBLE_PRECISION,
    ufo.KICK | ufo.DOUBLE_PRECISION | ufo.EXACT,
    ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED,
    ufo.DOUBLE_PRECISION,
    ufo.DOUBLE_PRECISION | ufo.ACHROMATIC,
    ufo.KICK | ufo.DOUBLE_PRECISION | ufo.LINEAR,
])
def test_every_element_kind(flags):
    """Every element class, misalignments, multipole errors, pole faces,
    fringe, wire, aperture, with observation points inside elements."""
    line = _all_elements_line()
    rng = np.random.default_rng(7)
    n = 128
    pool = np.stack([rng.normal(0, 2e-3, n), rng.normal(0, 2e-4, n), rng.normal(0, 2e-3, n),
                     rng.normal(0, 2e-4, n), rng.normal(0, 1e-3, n)], 1)
    where = [0.5, 2, 2.25, 4.5, 11, -1]
    gpu, strict, fast, moved = _track_both(line, pool, flags, 40, COORD6, where, dp=2e-3)
    r = compare_tracks(gpu, strict, fast, moved)
    print(f"flags {flags}: " + summary(r))
    assert_parity(r)


@pytest.mark.parametrize("flags", [ufo.KICK, 0, ufo.KICK | ufo.EXACT, ufo.KICK | ufo.FIVED])
def test_every_element_kind_fp32(flags):
    """The same line in the reference's default precision (FP32): no
    particle is regular at the 3e-13 level in FP32, so the bar is identical
    lost/alive flags and HIP's median and 90th-percentile error against the
    strict oracle within twice the reference's own (fast-math and one-ulp
    FP32 moves)."""
    line = _all_elements_line()
    rng = np.random.default_rng(7)
    n = 128
    pool = np.stack([rng.normal(0, 2e-3, n), rng.normal(0, 2e-4, n), rng.normal(0, 2e-3, n),
                     rng.normal(0, 2e-4, n), rng.normal(0, 1e-3, n)], 1)
    where = [0.5, 2, 2.25, 4.5, 11, -1]
    gpu, strict, fast, moved = _track_both(line, pool, flags, 40, COORD6, where, dp=2e-3)
    r = compare_tracks(gpu, strict, fast, moved)
    print(f"flags {flags}: " + summary(r))
    assert_parity(r, min_regular_frac=0.0)


def test_lattice_parameters_bound_per_particle():
    """Lattice values as per-particle parameters (k1 of a family, a dkN
    error, a displacement, a multipole kN, a length), tracked in one launch."""
    ring = lattices.fodo_ring(8)
    prm = ["x", "y", ("QF", "k1"), ("QD", "dk2"), ("SF", "k2"), ("QF", "dx"), (2, "length")]
    n = 96
    rng = np.random.default_rng(3)
    pool = np.stack([rng.normal(0, 1e-3, n), rng.normal(0, 1e-3, n), 1.2 + rng.normal(0, 0.02, n),
                     rng.normal(0, 0.5, n), 0.8 + rng.normal(0, 0.1, n), rng.normal(0, 1e-4, n),
                     2.0 + rng.normal(0, 0.01, n)], 1)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    gpu, strict, fast, moved = _track_both(ring, pool, flags, 100, prm, [-1])
    r = compare_tracks(gpu, strict, fast, moved)
    print(summary(r))
    assert_parity(r)


def test_fp32_thick_madx_vector():
    """The reference's default precision (FP32, thick): MAD-X vector of
    tests/tracking.py within its 1e-6 tolerance."""
    lat = ufo.Lattice(os.path.join(OPTICS_DIR, "fodo.mad"))
    tr = ufo.Track(lat.RING, turns=10, particles=2, where=[-1], flags=ufo.FIVED, parameters=["x"])
    tr.parameters[:, 0] = 0.001
    tr.run()
    x = np.array([-0.0008304642239, -0.001747306891, -0.0007418783837, 0.001079715459, 0.001713478007,
                  0.000462185744, -0.00129757264, -0.001629827644, -0.0001690544829, 0.001477701296])
    assert np.max(np.abs(tr.tracks[0][:, 0] - x)) < 1e-6
    assert np.max(np.abs(tr.tracks[1][:, 0] - x)) < 1e-6


def test_optics_alba_tunes():
    """ALBA periodic tunes and Twiss vs MAD-X (tests/optics.py) and vs the
    oracle's Optics."""
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    where = alba.RING.find(lambda e: isinstance(e, ufo.Quadrupole))[:20]
    fl = ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION
    o = ufo.Optics(alba.RING, where=list(where), flags=fl)
    o.run()
    assert abs(o.qx - 0.15543311) < 1e-8 and abs(o.qy - 0.362008337) < 1e-8
    r = runner.Optics(alba.RING, where=list(where), flags=fl)
    r.run()
    for a in ("bx", "by", "ax", "ay", "dx", "dpx", "mux", "muy"):
        np.testing.assert_allclose(getattr(o, a), getattr(r, a), rtol=1e-9, atol=1e-12)



def test_optics_alba_default_fp32():
    """tests/optics.py, chroma.py and radiation.py run Optics with its
    default flags (LINEAR | ACHROMATIC, FP32).  HIP in FP32: tunes within
    the reference test's 1e-5 of MAD-X; Twiss functions and chromaticity
    within twice the reference's own FP32 spread (its strict against its
    fast-math build) of the strict oracle, plus 1e-6."""
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    where = alba.RING.find(lambda e: hasattr(e, "k1") or hasattr(e, "k2"))
    o = ufo.Optics(alba.RING, where=list(where))
    o.run(threads=1)
    assert o.fp32
    assert abs(o.qx - 0.15543311) < 1e-5 and abs(o.qy - 0.362008337) < 1e-5
    strict = runner.Optics(alba.RING, where=list(where), flags=ufo.LINEAR | ufo.ACHROMATIC)
    fast = runner.Optics(alba.RING, where=list(where), flags=ufo.LINEAR | ufo.ACHROMATIC, mode="fast")
    strict.run()
    fast.run()
    for a in ("bx", "by", "ax", "ay", "dx", "dpx"):
        hip, ref, alt = (np.asarray(getattr(x, a), dtype=np.float64) for x in (o, strict, fast))
        scale = max(1.0, float(np.max(np.abs(ref))))
        spread = float(np.max(np.abs(alt - ref))) / scale
        err = float(np.max(np.abs(hip - ref))) / scale
        print(f"{a}: hip {err:.2e}, reference spread {spread:.2e}")
        assert err <= 2.0 * spread + 1e-6, a
    (n_hip, t_hip), (n_ref, t_ref), (n_alt, t_alt) = (ufo.chromaticity(x) for x in (o, strict, fast))
    for h, r, a in zip(n_hip + t_hip, n_ref + t_ref, n_alt + t_alt):
        assert abs(h - r) <= 2.0 * abs(a - r) + 1e-6 * max(1.0, abs(r))


@pytest.mark.parametrize("double", [True, False])
def test_optics_and_track_inside_thick_elements(double):
    """Fractional observation points (ufo/lattice.py:86-124 fractures an
    element at idx + f) inside ALBA's thick quadrupoles, combined-function
    bends and sextupoles: the HIP Optics equals the oracle's at every such
    point (FP64 1e-9; FP32 within twice the reference's own FP32 spread plus
    1e-6), and Track samples there agree with the oracle by the parity
    measure."""
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    ring = alba.RING
    pick = (ring.find(lambda e: isinstance(e, ufo.Quadrupole))[:6] + ring.find(lambda e: isinstance(e, ufo.Sbend))[:4]
            + ring.find(lambda e: isinstance(e, ufo.Sextupole))[:4])
    where = sorted(float(i + f) for i in pick for f in (0.25, 0.5, 0.8))
    fl = ufo.LINEAR | ufo.ACHROMATIC | (ufo.DOUBLE_PRECISION if double else 0)
    o = ufo.Optics(ring, where=list(where), flags=fl)
    o.run(threads=1)
    strict = runner.Optics(ring, where=list(where), flags=fl)
    strict.run()
    alt = runner.Optics(ring, where=list(where), flags=fl, mode="fast")
    alt.run()
    assert o.count == strict.count == len(where) + 1
    for a in ("bx", "by", "ax", "ay", "dx", "dpx"):
        hip, ref, fast = (np.asarray(getattr(x, a), dtype=np.float64) for x in (o, strict, alt))
        scale = max(1.0, float(np.max(np.abs(ref))))
        err = float(np.max(np.abs(hip - ref))) / scale
        bar = 1e-9 if double else 2.0 * float(np.max(np.abs(fast - ref))) / scale + 1e-6
        assert err <= bar, (a, err, bar)
    if not double:
        return
    tflags = ufo.DOUBLE_PRECISION
    bunch = lattices.gaussian_bunch(128, seed=21)
    bunch[:, [0, 2]] *= 0.5
    tr = ufo.Track(ring, flags=tflags, turns=20, particles=len(bunch), parameters=COORD6, where=list(where))
    tr.parameters[:] = bunch
    tr.run()
    runs = reference_runs(lambda pool, mode: runner.track(ring, pool, flags=tflags, turns=20, parameters=COORD6,
                                                          where=list(where), mode=mode), bunch)
    r = compare_tracks(tr.tracks, *runs)
    print(summary(r))
    # thick 6D with per-particle dp: the reference's own runs part at ~6e-13
    # (median) within 20 turns, so about a fifth of the particles stay under
    # the 3e-13 regularity bar (27/128 in the host build of this kernel)
    assert_parity(r, min_regular_frac=0.1)

def test_optics_consumers_alba():
    """chromaticity, emittance and rdt (ufo/chroma.py, ufo/radiation.py,
    ufo/rdt.py) on the HIP Optics of ALBA agree with the same functions on
    the oracle's Optics."""
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    fl = ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION
    cases = [(lambda e: hasattr(e, "k1") or hasattr(e, "k2"), ufo.chromaticity),
             (lambda e: hasattr(e, "angle"), ufo.emittance),
             (lambda e: isinstance(e, ufo.Sextupole), ufo.rdt)]
    for pick, fn in cases:
        where = list(alba.RING.find(pick))
        o = ufo.Optics(alba.RING, where=list(where), flags=fl)
        o.run()
        r = runner.Optics(alba.RING, where=list(where), flags=fl)
        r.run()
        a, b = fn(o), fn(r)
        if fn is ufo.emittance:
            a, b = (a.I1, a.I2, a.I4, a.I5, a.Ex), (b.I1, b.I2, b.I4, b.I5, b.Ex)
        else:
            a, b = np.ravel(a), np.ravel(b)
        print(fn.__name__, a)
        np.testing.assert_allclose(np.asarray(a, dtype=complex), np.asarray(b, dtype=complex), rtol=1e-9)


@pytest.mark.parametrize("flags", [ufo.DOUBLE_PRECISION, ufo.KICK | ufo.DOUBLE_PRECISION,
                                   ufo.KICK | ufo.DOUBLE_PRECISION | ufo.EXACT])
def test_elements_lattice(flags):
    """tests/elements.py's lattice (quadrupole, sbend with pole faces, drift,
    marker, sextupole, octupole multipole) tracked 5 passes with an
    observation point on the marker (never written, as in the reference)."""
    lat = ufo.Lattice(os.path.join(OPTICS_DIR, "elements.mad"))
    rng = np.random.default_rng(11)
    pool = np.stack([rng.normal(0, 1e-3, 64), rng.normal(0, 1e-4, 64), rng.normal(0, 1e-3, 64),
                     rng.normal(0, 1e-4, 64), rng.normal(0, 1e-3, 64)], 1)
    gpu, strict, fast, moved = _track_both(lat.TEST, pool, flags, 5, COORD6, [1, 3.5, -1])
    r = compare_tracks(gpu, strict, fast, moved)
    print(f"flags {flags}: " + summary(r))
    assert_parity(r)
    assert np.all(np.isnan(gpu[:, 10:, :]))          # samples 10..14 are never written


@pytest.mark.parametrize("double", [False, True])
def test_octupole_ring_closed_form(double):
    """tests/oct.py: oct.mad, FIVED, x0 = 0.01, 10 turns, against the exact
    map (mpmath, 200 bits)."""
    from test_reference_cases import _oct_closed_form
    lat = ufo.Lattice(os.path.join(OPTICS_DIR, "oct.mad"))
    flags = ufo.FIVED | (ufo.DOUBLE_PRECISION if double else 0)
    tr = ufo.Track(lat.RING, turns=10, particles=2, where=[-1], flags=flags, parameters=["x"])
    tr.parameters[:, 0] = 0.01
    tr.run(threads=1)
    exact = _oct_closed_form(0.01, 10, 10765.5878, 0.2876, 200)
    tol = 1e-13 if double else 2e-5
    np.testing.assert_allclose(tr.tracks[0][:, 0], exact[:, 0], rtol=tol, atol=tol * 0.01)
    np.testing.assert_allclose(tr.tracks[0][:, 1], exact[:, 1], rtol=tol, atol=tol * 1e-4)


def test_optics_periodic_equals_propagated():
    """tests/optics2.py on the HIP Optics: the propagated solution started
    from the periodic end values returns them."""
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    fl = ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION
    per = ufo.Optics(alba.RING, where=[-1], periodic=True, flags=fl)
    per.run(threads=1)
    pro = ufo.Optics(alba.RING, where=[-1], periodic=False, flags=fl)
    for a in ("ax", "ay", "bx", "by", "dx", "dpx"):
        setattr(pro, a + "0", getattr(per, a)[-1])
    pro.run(threads=1)
    for a in ("ax", "ay", "bx", "by", "dx", "dpx"):
        assert abs(getattr(per, a)[-1] - getattr(pro, a)[-1]) < 1e-9 * max(1.0, abs(getattr(per, a)[-1]))


def test_optics_chromaticity_by_dp_parameter():
    """tests/optics3.py on the HIP Optics: dp as a kernel parameter, LINEAR
    (chromatic, no sextupoles), tune shift over dp = 1e-4 against MAD-X's
    natural chromaticity (-40.13, -27.05) within 0.1, and equal to the
    oracle's finite difference to 1e-6."""
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    fl = ufo.LINEAR | ufo.DOUBLE_PRECISION
    res = []
    for make in (lambda: ufo.Optics(alba.RING, where=[], parameters=["dp"], flags=fl),
                 lambda: runner.Optics(alba.RING, where=[], parameters=["dp"], flags=fl)):
        o = make()
        o.parameters[0] = 0.0
        o.run() if isinstance(o, runner.Optics) else o.run(threads=1)
        q0 = (o.qx, o.qy)
        o.parameters[0] = 1e-4
        o.run() if isinstance(o, runner.Optics) else o.run(threads=1)
        res.append(((o.qx - q0[0]) / 1e-4, (o.qy - q0[1]) / 1e-4))
    print("dQ/ddp hip", res[0], "oracle", res[1])
    assert abs(res[0][0] + 40.13060686) < 0.1 and abs(res[0][1] + 27.05066939) < 0.1
    assert np.allclose(res[0], res[1], rtol=1e-6)


def test_sa_threads_and_refill():
    """Results do not depend on the number of resident lanes (persistent
    refill) nor on the order in which lanes claim particles — run with 64
    lanes and with the full device, claiming by amplitude (the default) and
    in pool order."""
    import torch
    ring = lattices.fodo_ring(16)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    pool = lattices.da_grid(40, 40, xlim=0.05, ymax=0.05)
    sa = ufo.StableAperture(ring, flags=flags, turns=300, particles=len(pool), parameters=["x", "y"])
    sa.parameters[:] = pool
    sa.run()
    full = sa.lost.copy()
    order = sa.order_device()
    assert order is not None and sorted(order.cpu().tolist()) == list(range(len(pool)))
    for claim in ("amplitude", None):
        sa.claim_order = claim
        sa._pool_dev.copy_(sa._pool_host)
        sa._lost_dev.fill_(-1)
        sa.launch(lanes=64)
        sa._lost_host.copy_(sa._lost_dev)
        torch.cuda.synchronize()
        assert np.array_equal(sa.lost, full), claim
    ref = runner.stable_aperture(ring, pool, flags=flags, turns=300, parameters=["x", "y"])
    assert np.array_equal(full, ref)
    assert 0 < np.sum(full < 300) < len(pool)


@pytest.mark.parametrize("double", [False, True])
def test_alba_stable_aperture_howto_grid(double):
    """HOWTO.md:332-370 12x12 ALBA grid, 1000 turns: in FP64 the HIP kernel
    gives the oracle's loss turns on every particle whose loss turn is
    regular (strict, fast-math and one-ulp-moved oracle runs agree); in FP32 (the documented run) every cell documented as 0, and
    every 1000 cell whose neighbours are all 1000, is reproduced (FP32
    escapes next to the boundary differ between any two evaluations)."""
    import json
    g = json.load(open(os.path.join(os.path.dirname(OPTICS_DIR), "reference_values.json")))["alba_sa_grid_fp32"]
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    n = 12
    x = np.linspace(-0.04, 0.04, n)
    xx, yy = np.meshgrid(x, x)
    pool = np.stack([xx.ravel(), yy.ravel()], 1)
    flags = ufo.FIVED | (ufo.DOUBLE_PRECISION if double else 0)
    sa = ufo.StableAperture(alba.RING, flags=flags, turns=1000, particles=n * n, parameters=["x", "y"])
    sa.parameters[:] = pool
    sa.run()
    strict, regular = loss_turn_runs(
        lambda p, m: runner.stable_aperture(alba.RING, p, flags=flags, turns=1000, parameters=["x", "y"], mode=m),
        pool, fp32=not double)
    print(f"regular {regular.sum()}/{n * n}; gpu != strict on {np.sum(sa.lost != strict)}")
    if double:
        assert np.array_equal(sa.lost[regular], strict[regular])
    else:
        # FP32: chaotic escapes near the boundary differ by a turn between
        # any two FP32 evaluations; the documented 0/1000 cells must agree
        ref = np.array(g["lost"])
        pad = np.pad(ref, 1, mode="edge")
        neigh = [pad[1:-1, :-2], pad[1:-1, 2:], pad[:-2, 1:-1], pad[2:, 1:-1]]
        interior = (ref == 0) | ((ref == 1000) & np.all([nb == 1000 for nb in neigh], axis=0))
        assert np.all(sa.lost[interior.ravel()] == ref.ravel()[interior.ravel()])


def test_reference_aperture_case():
    """tests/aperture.py on the HIP Track: one Aperture('ap', window='x > 0.1'),
    FIVED, FP32 (the reference test's flags), x = 0 and x = 0.2.  The window
    is spliced as ``if (x > 0.1 == 0) x = NAN;`` (ufo/elements.py:518), so x = 0
    is lost and x = 0.2 kept, as in the oracle; a two-term window binds
    ``a && b == 0`` as ``a && (b == 0)``."""
    line = ufo.Line("a_line")
    line.append(ufo.Aperture("ap", window="x > 0.1"))
    tr = ufo.Track(line, turns=1, particles=2, where=[-1], flags=ufo.FIVED, parameters=["x"])
    tr.parameters[0, 0] = 0.0
    tr.parameters[1, 0] = 0.2
    tr.run(threads=1)
    ref = runner.track(line, tr.parameters.copy(), flags=ufo.FIVED, turns=1, parameters=["x"], where=[-1])
    assert np.isnan(tr.tracks[0, 0, 0]) and tr.tracks[1, 0, 0] == np.float32(0.2)
    np.testing.assert_array_equal(tr.tracks[:, :, :4], ref[:, :, :4])

    line = ufo.Line("l", [ufo.Aperture("ap", window="x > 0.1 && y > 0.1")])
    pool = np.array([[0.2, 0.0], [0.2, 0.2], [0.0, 0.0]])
    fl = ufo.FIVED | ufo.DOUBLE_PRECISION
    tr = ufo.Track(line, turns=1, particles=3, where=[-1], flags=fl, parameters=["x", "y"])
    tr.parameters[:] = pool
    tr.run()
    assert np.isnan(tr.tracks[0, 0, 0]) and tr.tracks[1, 0, 0] == 0.2 and tr.tracks[2, 0, 0] == 0.0


def test_reference_align_case():
    """tests/align.py on the HIP Optics: the FODO with QF displaced by
    dx = 1e-4 has the dispersion of the line whose QF is a combined bend of
    angle 1e-4 (optics/align.mad), within 1e-6, in the reference test's FP32
    and in FP64; each equals the oracle's Optics."""
    for fl, tol in [(ufo.LINEAR | ufo.ACHROMATIC, 1e-6),
                    (ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION, 1e-12)]:
        lat = ufo.Lattice(os.path.join(OPTICS_DIR, "align.mad"))
        lat.QF.dx = 1e-4
        fodo = ufo.Optics(lat.FODO, where=[-1], flags=fl)
        equivalent = ufo.Optics(lat.EQUIVALENT, where=[-1], flags=fl)
        fodo.run(threads=1)
        equivalent.run(threads=1)
        assert abs(fodo.dx[0] - equivalent.dx[0]) < 1e-6
        for o, line in [(fodo, lat.FODO), (equivalent, lat.EQUIVALENT)]:
            r = runner.Optics(line, where=[-1], flags=fl)
            r.run()
            assert abs(o.dx[0] - r.dx[0]) <= tol * max(1.0, abs(r.dx[0]))
            assert abs(o.qx - r.qx) <= tol and abs(o.qy - r.qy) <= tol


def test_reference_thick_vs_kick_case():
    """tests/thick_vs_kick.py on the HIP Optics: ALBA with 16 quadrupole and
    48 bend slices (set through ufo.DEFAULT_QUADRUPOLE_SLICES /
    DEFAULT_BEND_SLICES before the lattice is loaded), teapot and thick tunes
    agree within 5e-4; in FP64 each tune equals the oracle's to 1e-10."""
    saved = ufo.DEFAULT_QUADRUPOLE_SLICES, ufo.DEFAULT_BEND_SLICES
    try:
        ufo.DEFAULT_QUADRUPOLE_SLICES = 16
        ufo.DEFAULT_BEND_SLICES = 48
        alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    finally:
        ufo.DEFAULT_QUADRUPOLE_SLICES, ufo.DEFAULT_BEND_SLICES = saved
    assert all(e.slices == 16 for e in alba.RING.flatten() if isinstance(e, ufo.Quadrupole))
    for double in (False, True):
        fl = ufo.LINEAR | ufo.ACHROMATIC | (ufo.DOUBLE_PRECISION if double else 0)
        kick = ufo.Optics(alba.RING, flags=fl | ufo.KICK)
        thick = ufo.Optics(alba.RING, flags=fl)
        kick.run(threads=1)
        thick.run(threads=1)
        assert abs(thick.qx - kick.qx) < 5e-4 and abs(thick.qy - kick.qy) < 5e-4
        if double:
            for o, f in [(kick, fl | ufo.KICK), (thick, fl)]:
                r = runner.Optics(alba.RING, flags=f)
                r.run()
                assert abs(o.qx - r.qx) < 1e-10 and abs(o.qy - r.qy) < 1e-10


def test_nan_start_and_no_parameters():
    """IEEE semantics on the device: a particle starting with x = NaN is lost
    at turn 0 in StableAperture (fabs(NaN) > 1 is false, isnan catches it)
    and its Track samples are NaN, as in the strict oracle; a Track with no
    parameters runs every particle from the origin."""
    ring = lattices.fodo_ring(4)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    pool = lattices.gaussian_bunch(64, seed=2)
    pool[[3, 17], 0] = np.nan
    sa = ufo.StableAperture(ring, flags=flags, turns=50, particles=len(pool), parameters=COORD6)
    sa.parameters[:] = pool
    sa.run()
    ref = runner.stable_aperture(ring, pool, flags=flags, turns=50, parameters=COORD6)
    assert sa.lost[3] == 0 and sa.lost[17] == 0
    assert np.array_equal(sa.lost, ref)
    gpu, strict, fast, moved = _track_both(ring, pool, flags, 20, COORD6, [-1])
    assert np.array_equal(np.isnan(gpu[:, :, 0]), np.isnan(strict[:, :, 0]))
    assert np.all(np.isnan(gpu[[3, 17], :, 0]))
    tr = ufo.Track(ring, flags=flags, turns=5, particles=8, parameters=[], where=[-1])
    tr.run()
    assert np.all(tr.tracks == 0.0)


@pytest.mark.parametrize("flags", [ufo.KICK | ufo.DOUBLE_PRECISION, ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED])
def test_observation_point_in_every_cell(flags):
    """One observation point inside each cell's first quadrupole of the
    16-cell ring: the stores sit in the period loop; samples in lattice
    order, parity with the oracle (and, FIVED, the host build bit for bit)."""
    ring = lattices.fodo_ring(16)
    cell = len(ring.flatten()) // 16
    where = [c * cell + 0.5 for c in range(16)] + [-1]
    bunch = lattices.gaussian_bunch(256, seed=8)
    gpu, strict, fast, moved = _track_both(ring, bunch, flags, 50, COORD6, where)
    assert gpu.shape == (256, 50 * 17, 6)
    r = compare_tracks(gpu, strict, fast, moved)
    print(summary(r))
    assert_parity(r)


@pytest.mark.parametrize("double", [True, False])
def test_howto_time_dependent_kicker(double):
    """HOWTO.md:86-120 on the GPU: a user subclass of Multipole whose kick
    reads the turn counter, inserted at the start of ALBA (thick, FIVED),
    100 particles of the equilibrium emittance, 65 turns: the emittance
    jumps after turn 20, and the tracks agree with the oracle (FP64: the
    parity bar; FP32, the HOWTO's own precision: the statistical bar)."""
    from test_reference_cases import nlk_ring, nlk_bunch
    ring = nlk_ring()
    pool = nlk_bunch()
    flags = ufo.FIVED | (ufo.DOUBLE_PRECISION if double else 0)
    gpu, strict, fast, moved = _track_both(ring, pool, flags, 65, ["x", "px"], [-1])
    r = compare_tracks(gpu, strict, fast, moved)
    print(summary(r))
    assert_parity(r, min_regular_frac=0.0)
    ex = np.mean(gpu[:, :, 0] ** 2 / 11.26 + gpu[:, :, 1] ** 2 * 11.26, axis=0)
    assert abs(ex[19] - ex[0]) < 0.15 * ex[0] and ex[25] > 1.1 * ex[19]


def test_alba_thick_small_amplitude_fp64():
    """Thick ALBA (FIVED, FP64) with a bunch of the equilibrium emittance,
    200 turns: nearly linear motion, where the reference's own rounding
    noise is smallest and a one-ulp bias in a folded matrix coefficient
    would dominate (the constants are folded correctly rounded)."""
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    rng = np.random.default_rng(5)
    n = 128
    pool = np.stack([rng.normal(0, 1.5e-4, n), rng.normal(0, 1.5e-5, n),
                     rng.normal(0, 5e-5, n), rng.normal(0, 5e-6, n)], 1)
    flags = ufo.FIVED | ufo.DOUBLE_PRECISION
    gpu, strict, fast, moved = _track_both(alba.RING, pool, flags, 200, ["x", "px", "y", "py"], [-1])
    r = compare_tracks(gpu, strict, fast, moved)
    print(summary(r))
    assert_parity(r, min_regular_frac=0.0)


@pytest.mark.parametrize("amp,turns", [(1.0, 200), (10.0, 100)])
def test_alba_thick_6d_small_amplitude(amp, turns):
    """Thick ALBA in 6D (dp per particle, so every thick coefficient is a
    per-particle invariant: correctly rounded sin/cos/sinh/cosh, products and
    quotients carried with their exact low parts), bunches near the
    equilibrium emittance: parity with the oracle, and bit for bit the host
    build of the same kernel."""
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    rng = np.random.default_rng(5)
    n = 128
    pool = np.stack([rng.normal(0, 1.5e-4 * amp, n), rng.normal(0, 1.5e-5 * amp, n), rng.normal(0, 5e-5 * amp, n),
                     rng.normal(0, 5e-6 * amp, n), rng.normal(0, 1e-3, n)], 1)
    gpu, strict, fast, moved = _track_both(alba.RING, pool, ufo.DOUBLE_PRECISION, turns, COORD6, [-1])
    r = compare_tracks(gpu, strict, fast, moved)
    print(summary(r))
    assert_parity(r, min_regular_frac=0.0)


GOLDEN_LINES = [("align", "FODO"), ("align", "EQUIVALENT"), ("bend", "SRING"), ("bend", "RRING"), ("drift", "RING"),
                ("elements", "TEST"), ("fodo", "RING"), ("fodo2", "RING"), ("multipole", "FODO"), ("oct", "RING"),
                ("triplet", "TRIPLET"), ("alba", "RING")]


@pytest.mark.parametrize("flags", [ufo.DOUBLE_PRECISION, ufo.KICK | ufo.DOUBLE_PRECISION,
                                   ufo.KICK | ufo.EXACT | ufo.DOUBLE_PRECISION, ufo.FIVED])
@pytest.mark.parametrize("lattice,line", GOLDEN_LINES)
def test_golden_lattices(lattice, line, flags):
    """Every line of every lattice file the reference ships (optics/*.mad,
    copied as fixtures) tracked 30 turns in 6D thick, 6D teapot, 6D teapot
    with exact drifts (FP64) and FIVED thick (FP32, the reference's default
    precision): parity with the oracle."""
    lat = ufo.Lattice(os.path.join(OPTICS_DIR, f"{lattice}.mad"))
    ring = getattr(lat, line)
    import zlib
    rng = np.random.default_rng(zlib.crc32(f"{lattice}.{line}".encode()))
    n = 64
    pool = np.stack([rng.normal(0, 5e-4, n), rng.normal(0, 5e-5, n), rng.normal(0, 5e-4, n),
                     rng.normal(0, 5e-5, n), rng.normal(0, 5e-4, n)], 1)
    gpu, strict, fast, moved = _track_both(ring, pool, flags, 30, COORD6, [-1])
    r = compare_tracks(gpu, strict, fast, moved)
    print(f"{lattice}.{line} flags {flags}: " + summary(r))
    # strong sextupoles and octupoles make every particle of some of these
    # lines chaotic within 30 turns: the per-particle factor 8 (parity.py)
    assert_parity(r, min_regular_frac=0.0, tail_factor=8.0)


def test_empty_inputs_on_every_entry_point():
    """No particles: every entry point runs without a launch and returns
    empty results of the documented shapes (the reference cannot size an
    OpenCL buffer of zero bytes, ufo/track.py:57); an AdjointTracker with no
    particles gives a zero gradient."""
    ring = lattices.fodo_ring(4)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    tr = ufo.Track(ring, flags=flags, turns=5, particles=0, parameters=COORD6, where=[-1])
    tr.run()
    assert tr.tracks.shape == (0, 5, 6)
    sa = ufo.StableAperture(ring, flags=flags, turns=5, particles=0, parameters=COORD6)
    sa.run()
    assert sa.lost.shape == (0,)
    co = ufo.ClosedOrbit(ring, flags=flags, particles=0, parameters=["x"])
    co.run()
    assert co.orbits.shape == (0, 4) and co.converged.shape == (0,)
    tm = ufo.TangentMap(ring, flags=flags, turns=3, particles=0, parameters=COORD6)
    tm.run()
    assert tm.state.shape == (0, 4) and tm.jacobian.shape == (0, 4, 4)
    shared = [("QF", "k1")]
    at = ufo.AdjointTracker(ring, flags, 10, COORD6, shared)
    x0 = torch.zeros((0, 5), dtype=torch.float64, device=at.device)
    xf, lost = at.forward(x0, torch.tensor([1.2], dtype=torch.float64))
    assert xf.shape == (0, 4) and lost.shape == (0,)
    gx0, g = at.backward(torch.zeros((0, 4), dtype=torch.float64, device=at.device))
    assert gx0.shape == (0, 4) and g.shape == (1,) and float(g[0]) == 0.0
