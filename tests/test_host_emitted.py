"""The emitted Track kernel, compiled for the host with g++ (same IEEE
operations in the same order as on the GPU, see host_emulation.py), against
the oracle on CPU: the generated code's parity with the reference semantics
is checked here without a GPU, on the teapot lattices where the host build is
bit-exact with the device."""
import numpy as np
import pytest

import ufo_amd as ufo
from ufo_amd import codegen, lattices
from oracle import runner

import host_emulation
from parity import assert_parity, compare_tracks, reference_runs, summary

COORD6 = ["x", "px", "y", "py", "dp"]


def _host_vs_oracle(line, pool, flags, turns, parameters, where, dp=0.0):
    k = codegen.track_kernel(line, flags, turns, parameters, where, dp)
    samples = turns * max(k.dumps_per_turn, 1)
    host = host_emulation.run_track(k, pool, samples)

    def oracle(p, mode):
        return runner.track(line, p, flags=flags, turns=turns, parameters=parameters, where=where, dp=dp, mode=mode)

    return (host,) + reference_runs(oracle, pool)


def test_fodo_6d_host_build_matches_oracle():
    ring = lattices.fodo_ring(16)
    bunch = lattices.gaussian_bunch(128, seed=5)
    bunch[:4, 0] = [0.035, 0.045, 0.05, 0.06]          # some lost at the aperture
    host, strict, fast, moved = _host_vs_oracle(ring, bunch, ufo.KICK | ufo.DOUBLE_PRECISION, 60, COORD6, [-1])
    r = compare_tracks(host, strict, fast, moved)
    print(summary(r))
    assert_parity(r, 0.3)
    assert np.array_equal(np.isnan(host[:, :, 0]), np.isnan(strict[:, :, 0]))


@pytest.mark.parametrize("flags", [ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED,
                                   ufo.KICK | ufo.DOUBLE_PRECISION | ufo.LINEAR])
def test_dba_period_host_build_matches_oracle(flags):
    """Two DBA periods with misaligned magnets and multipole errors bound as
    per-particle parameters, observation points inside elements."""
    line = ufo.Line("L", [lattices.dba_period("A"), lattices.dba_period("B")])
    prm = ["x", "y", ("QF2A", "dx"), ("SF1B", "dk3"), ("QD1A", "dk2s")]
    rng = np.random.default_rng(2)
    n = 96
    pool = np.stack([rng.normal(0, 3e-3, n), rng.normal(0, 2e-3, n), rng.normal(0, 5e-5, n),
                     rng.normal(0, 50.0, n), rng.normal(0, 0.5, n)], 1)
    host, strict, fast, moved = _host_vs_oracle(line, pool, flags, 40, prm, [3.5, 17, -1])
    r = compare_tracks(host, strict, fast, moved)
    print(summary(r))
    assert_parity(r)


@pytest.mark.parametrize("splits", [[1], [7, 8, 31], list(range(1, 60))])
def test_track_split_over_launches_is_bit_identical(splits):
    """Track.run splits the turns over launches that resume from the saved
    (x, px, y, py); on the host build of the same kernel any split gives the
    one-launch result bit for bit: particles lost at the aperture part-way
    (the kernel writes NaN into their later samples), 6D with dp per
    particle, and a fractional observation point on a thin element that is
    never written (sample numbering turn * dumps_per_turn + k, the tail left
    unwritten by the kernel)."""
    ring = lattices.fodo_ring(16)
    bunch = lattices.gaussian_bunch(100, seed=9)
    bunch[:6, 0] = [0.035, 0.038, 0.045, 0.05, 0.0399, -0.039]
    sf = ring.find(lambda e: isinstance(e, ufo.Multipole))[0]
    where = [2.5, sf + 0.5, -1]
    k = codegen.track_kernel(ring, ufo.KICK | ufo.DOUBLE_PRECISION, 60, COORD6, where, 0.0)
    assert k.dumps_per_turn == 2 and k.turns == 60
    samples = 60 * len(where)
    one = host_emulation.run_track(k, bunch, samples)
    split = host_emulation.run_track(k, bunch, samples, splits=splits)
    lost = np.isnan(one[:, 119, 0])              # the last written sample
    assert 0 < lost.sum() < len(bunch) and np.all(np.isnan(one[:, 120:, :]))
    np.testing.assert_array_equal(split.view(np.uint64), one.view(np.uint64))
    # the kernel itself writes NaN into every sample of its turns after a
    # loss (the driver fills only the never-written tail): from a buffer of
    # 7.0, every one of the 120 numbered samples is written, the tail is not
    for cut in ((), splits):
        mark = host_emulation.run_track(k, bunch, samples, splits=cut, fill=7.0)
        np.testing.assert_array_equal(mark[:, :120].view(np.uint64), one[:, :120].view(np.uint64))
        assert np.all(mark[:, 120:] == 7.0)


def test_observation_point_in_every_cell_of_a_periodic_ring():
    """An observation point inside the first quadrupole of each of the 16
    cells: the repeated cell (with its store) becomes the period loop, and
    the samples are numbered in lattice order as in the reference."""
    ring = lattices.fodo_ring(16)
    cell = len(ring.flatten()) // 16
    where = [c * cell + 0.5 for c in range(16)] + [-1]
    k = codegen.track_kernel(ring, ufo.KICK | ufo.DOUBLE_PRECISION, 8, COORD6, where, 0.0)
    assert k.dumps_per_turn == 17 and k.period[2] >= 2 and "for (int rep" in k.src
    bunch = lattices.gaussian_bunch(64, seed=8)
    host, strict, fast, moved = _host_vs_oracle(ring, bunch, ufo.KICK | ufo.DOUBLE_PRECISION, 8, COORD6, where)
    r = compare_tracks(host, strict, fast, moved)
    print(summary(r))
    assert_parity(r)


@pytest.mark.parametrize("flags,params,where,turns,splits,n,scale", [
    (ufo.KICK | ufo.DOUBLE_PRECISION, ["x", "px", "y", "py", "dp"], [-1], 20, [], 77, 1e-4),
    (ufo.KICK | ufo.DOUBLE_PRECISION, ["x", "px", "y", "py", "dp"], [0, 3, 9, -1], 20, [7, 13], 300, 1e-4),
    (ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED, ["x", "y"], [2, -1], 9, [1, 8], 65, 1e-4),
    (ufo.KICK | ufo.EXACT | ufo.DOUBLE_PRECISION, ["x", "px", "y", "py", "dp"], [-1], 10, [5], 1, 1e-4),
    (ufo.KICK, ["x", "px", "y", "py", "dp"], [1, -1], 10, [3], 129, 1e-4),
    (ufo.DOUBLE_PRECISION, ["x", "px", "y", "py", "dp"], [-1, 4], 12, [6], 64, 1e-4),
    (ufo.KICK | ufo.DOUBLE_PRECISION, [], [-1], 5, [], 4, 1e-4),
    (ufo.KICK | ufo.DOUBLE_PRECISION, ["x", "px", "y", "py", "dp"], [0, 3, -1], 20, [7, 13], 300, 2e-4),
    (ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED, ["x", "y"], [2, -1], 9, [1, 8], 65, 8e-4),
])
def test_emitted_track_kernel_stays_in_bounds(flags, params, where, turns, splits, n, scale):
    """The emitted Track kernel and its resumable launches, built for the
    host under AddressSanitizer and UndefinedBehaviorSanitizer with every
    buffer exactly as large as the launch needs: particle counts that do
    not fill a block, several observation points, launches cut at arbitrary
    turns, FIVED, EXACT, thick, FP32 and no parameter columns; with larger
    pool values (scale 2e-4 in 6D: 168 of 300 particles lost over the turns;
    8e-4 in FIVED: 39 of 65, 13 in the first turn) the NaN samples after a
    loss are written too, also by launches that resume a lost particle."""
    k = codegen.track_kernel(lattices.fodo_ring(4), flags, turns, params, where, 0.0)
    rc, report = host_emulation.sanitized_run(k, n, len(params), splits, scale=scale)
    assert rc == 0, report[-2000:]



@pytest.mark.parametrize("kind,double", [("optics", True), ("optics", False), ("jvp", True), ("co", True),
                                         ("co", False)])
def test_emitted_optics_tangent_and_orbit_kernels_stay_in_bounds(kind, double):
    """The Optics (fractional and end-of-line observation points), JVP and
    ClosedOrbit kernels, built for the host under AddressSanitizer and
    UndefinedBehaviorSanitizer with exactly sized buffers and particle
    counts that do not fill a block."""
    from ufo_amd import tangent
    ring = lattices.fodo_ring(3)
    fl = ufo.DOUBLE_PRECISION if double else 0
    samples = 0
    if kind == "optics":
        k = codegen.optics_kernel(ring, ufo.LINEAR | ufo.ACHROMATIC | fl, [0.5, 2, 4.25, -1], [])
        samples = max(4, k.dumps_per_turn)
        cols, n = 1, 1
    elif kind == "jvp":
        cols, n = 5, 77
        k = tangent.jvp_kernel(ring, ufo.KICK | fl, 3, ["x", "px", "y", "py", "dp"], 0.0)
    else:
        cols, n = 2, 70
        k = tangent.co_kernel(ring, fl, [(1, "dx"), (3, "dy")], 8, 1e-15)
    rc, report = host_emulation.sanitized_run_kind(k, n, cols, samples)
    assert rc == 0, report[-2000:]


def _one_turn_residual(ring, flags, prm, pool, orbits):
    """max |F(X) - X| of the oracle's one-turn map at dp = 0 for each row."""
    full = np.concatenate([orbits, pool[:, 4:]], axis=1)
    f = runner.track(ring, full, flags=flags | ufo.FIVED, turns=1, parameters=prm, where=[-1])[:, 0, :4]
    return np.max(np.abs(f - orbits), axis=1)


def test_closed_orbit_keeps_the_best_evaluated_iterate():
    """The ClosedOrbit kernel built for the host: a converged orbit is a
    fixed point of the oracle's one-turn map and flagged converged; a guess
    lost in its first turn comes back unchanged and not converged; with one
    Newton step allowed the guess (the only evaluated iterate) comes back,
    never the unevaluated step; with tol = 0 (never reached) the returned
    orbit's residual does not grow as the step cap rises."""
    from ufo_amd import tangent
    ring = lattices.fodo_ring(16)
    fl = ufo.KICK | ufo.DOUBLE_PRECISION
    prm = ["x", "px", "y", "py", (0, "dx"), (3, "dy")]   # QF and QD of the first cell
    pool = np.array([[0.0, 0.0, 0.0, 0.0, 1e-4, -5e-5],
                     [0.05, 0.0, 0.0, 0.0, 1e-4, 0.0],            # beyond the 40 mm aperture: lost in turn 1
                     [2e-3, 1e-4, -1e-3, 2e-5, 1e-4, 5e-5]])
    orb, steps, conv = host_emulation.run_orbits(tangent.co_kernel(ring, fl, prm, 64, 1e-15), pool)
    assert conv.tolist() == [True, False, True]
    assert np.array_equal(orb[1], pool[1, :4]) and steps[1] == 0
    res = _one_turn_residual(ring, fl, prm, pool, orb)
    assert res[0] < 1e-15 and res[2] < 1e-15 and np.max(np.abs(orb[0])) > 1e-6
    one, steps1, conv1 = host_emulation.run_orbits(tangent.co_kernel(ring, fl, prm, 1, 1e-15), pool)
    assert np.array_equal(one, pool[:, :4]) and not conv1.any() and steps1.tolist() == [1, 0, 1]
    prev = None
    for cap in range(1, 7):
        o, s, c = host_emulation.run_orbits(tangent.co_kernel(ring, fl, prm, cap, 0.0), pool[2:])
        r = _one_turn_residual(ring, fl, prm, pool[2:], o)[0]
        assert not c[0] and s[0] == cap
        if prev is not None:
            assert r <= prev * (1 + 1e-6) + 1e-17, (cap, r, prev)
        prev = r
    assert prev < 1e-15


def test_seed_closed_orbit_lost_seed_is_not_converged():
    """The per-seed closed-orbit kernel built for the host: seeds from a
    zero guess converge to oracle fixed points; a seed whose guess is lost in
    its first turn is reported not converged and returned unchanged."""
    from ufo_amd import batch
    from ufo_amd.batch_adjoint import co_seed_kernel
    ring, prm, vals, _, _ = lattices.dba_family_model(seeds=3)
    fl = ufo.KICK | ufo.DOUBLE_PRECISION
    k, layout = co_seed_kernel(ring, fl, prm, 32, 1e-13)
    guess = np.zeros((3, 4))
    guess[1, 0] = 0.05                                            # beyond the 30 mm aperture
    orb, steps, conv = host_emulation.run_seed_orbits(k, batch.build_table(layout, vals), guess)
    assert conv.tolist() == [True, False, True]
    assert np.array_equal(orb[1], guess[1]) and steps[1] == 0
    ok = [0, 2]
    f = runner.track(ring, np.concatenate([orb[ok], vals[ok]], axis=1), flags=fl | ufo.FIVED, turns=1,
                     parameters=["x", "px", "y", "py"] + prm, where=[-1])[:, 0, :4]
    assert np.max(np.abs(f - orb[ok])) < 1e-13 and np.max(np.abs(orb[ok])) > 1e-7
