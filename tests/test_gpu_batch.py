"""GPU checks of the batched (multi-seed) dynamic-aperture tracker against
the oracle, where each particle carries its seed's error values as
per-particle parameters (the reference's way of batching lattice variants,
HOWTO.md:20-22)."""
import os

import numpy as np
import pytest

import ufo_amd as ufo
from ufo_amd import lattices
from ufo_amd.batch import BatchedAperture
from oracle import runner

from conftest import OPTICS_DIR
from parity import loss_turn_runs

pytestmark = pytest.mark.gpu


def _setup(seeds, count_range, turns, amplify=1.0):
    ring, prm, vals, counts = lattices.dba_error_model(periods=4, seeds=seeds, count_range=count_range)
    vals = vals * amplify
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    ba = BatchedAperture(ring, flags, turns, prm, vals, counts, block_turns=64)
    pool = np.concatenate([lattices.seed_grid(int(c), xlim=0.03, ymax=0.015) for c in counts])
    ba.parameters[:] = pool
    return ring, prm, vals, counts, flags, ba, pool


def test_batched_matches_oracle_per_particle_parameters():
    turns = 300
    ring, prm, vals, counts, flags, ba, pool = _setup(3, (150, 400), turns)
    ba.run()
    seed_of = np.repeat(np.arange(len(counts)), counts)
    full = np.concatenate([pool, vals[seed_of]], axis=1)
    strict, regular = loss_turn_runs(
        lambda p, m: runner.stable_aperture(ring, p, flags=flags, turns=turns, parameters=["x", "y"] + prm, mode=m),
        full)
    lost = ba.lost.astype(np.int64)
    print(f"particles {len(pool)}, lost {np.sum(strict < turns)}, regular {regular.sum()}, "
          f"differ {np.sum(lost != strict)}")
    assert 0.2 * len(pool) < np.sum(strict < turns) < 0.9 * len(pool)
    assert np.array_equal(lost[regular], strict[regular])


def test_batched_thick_alba_per_seed_quadrupoles():
    """Thick elements (no KICK): three variants of ALBA with per-seed QF1 and
    QD2 gradients (the thick matrices' square roots and trigonometric
    functions of the per-seed strengths live in the seed table), FIVED, 300
    turns over the HOWTO's 12 x 12 grid; loss turns equal the oracle's on
    every regular particle."""
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    flags = ufo.DOUBLE_PRECISION | ufo.FIVED
    prm = [("QF1", "k1"), ("QD2", "k1")]
    qf1 = [e.k1 for e in alba.RING.flatten() if e.label == "QF1"][0]
    qd2 = [e.k1 for e in alba.RING.flatten() if e.label == "QD2"][0]
    vals = np.array([[qf1, qd2], [qf1 * 1.01, qd2], [qf1, qd2 * 0.99]])
    xx, yy = np.meshgrid(np.linspace(-0.04, 0.04, 12), np.linspace(-0.04, 0.04, 12))
    grid = np.stack([xx.ravel(), yy.ravel()], 1)
    counts = np.array([len(grid)] * 3)
    turns = 300
    ba = BatchedAperture(alba.RING, flags, turns, prm, vals, counts, block_turns=64)
    pool = np.concatenate([grid] * 3)
    ba.parameters[:] = pool
    ba.run()
    seed_of = np.repeat(np.arange(3), counts)
    full = np.concatenate([pool, vals[seed_of]], axis=1)
    strict, regular = loss_turn_runs(
        lambda p, m: runner.stable_aperture(alba.RING, p, flags=flags, turns=turns, parameters=["x", "y"] + prm,
                                            mode=m),
        full)
    lost = ba.lost.astype(np.int64)
    print(f"lost {np.sum(strict < turns)} of {len(pool)}, regular {regular.sum()}, differ {np.sum(lost != strict)}")
    assert np.array_equal(lost[regular], strict[regular])
    assert 0 < np.sum(strict < turns) < len(pool) and regular.mean() > 0.8
    # the variants differ: the seeds' tables are used
    assert not np.array_equal(strict[:144], strict[144:288]) or not np.array_equal(strict[:144], strict[288:])


def test_batched_fp32_matches_plain_kernel_and_oracle():
    """The reference's default precision (FP32), three error seeds of the
    4-period ring, 200 turns.  The batched kernel, whose per-seed table is
    evaluated in FP32 on the host, gives exactly the loss turns of the plain
    StableAperture kernel with the errors as per-particle parameters (same
    coefficient values, same emitted arithmetic).  Against the oracle: in
    FP32 the loss turn of a particle near the boundary depends on rounding,
    so the bar is statistical — the GPU differs from the strict oracle on no
    more particles than twice the oracle's fast-math build does (+2)."""
    turns = 200
    ring, prm, vals, counts = lattices.dba_error_model(periods=4, seeds=3, count_range=(120, 300))
    flags = ufo.KICK | ufo.FIVED
    ba = BatchedAperture(ring, flags, turns, prm, vals, counts, block_turns=64)
    pool = np.concatenate([lattices.seed_grid(int(c), xlim=0.03, ymax=0.015) for c in counts])
    ba.parameters[:] = pool
    ba.run()
    seed_of = np.repeat(np.arange(len(counts)), counts)
    full = np.concatenate([pool, vals[seed_of]], axis=1)
    sa = ufo.StableAperture(ring, flags=flags, turns=turns, particles=len(full), parameters=["x", "y"] + prm)
    sa.parameters[:] = full
    sa.run()
    strict = runner.stable_aperture(ring, full, flags=flags, turns=turns, parameters=["x", "y"] + prm)
    fast = runner.stable_aperture(ring, full, flags=flags, turns=turns, parameters=["x", "y"] + prm, mode="fast")
    lost = ba.lost.astype(np.int64)
    d_gpu, d_ref = int(np.sum(lost != strict)), int(np.sum(fast != strict))
    print(f"FP32: lost {np.sum(strict < turns)} of {len(pool)}; differ from strict: batched {d_gpu}, "
          f"fast-math oracle {d_ref}; batched vs plain kernel {np.sum(lost != sa.lost)}")
    assert np.array_equal(lost, sa.lost.astype(np.int64))
    assert d_gpu <= 2 * d_ref + 2
    assert 0.2 * len(pool) < np.sum(strict < turns) < 0.9 * len(pool)


def test_graph_replay_and_launch_shape_do_not_change_results():
    turns = 256
    _, _, _, counts, _, ba, pool = _setup(5, (64, 700), turns)
    ba.run()
    first = ba.lost.copy()
    ba.run()                                # replays the captured graph
    assert np.array_equal(ba.lost, first)
    ba.block_turns = 32
    ba.run()                                # another block size: same physics
    assert np.array_equal(ba.lost, first)
    ba.particles_per_lane, ba.workgroup = 2, 64
    ba.run()                                # two particles per lane, 64-lane workgroups
    assert np.array_equal(ba.lost, first)
    ba.particles_per_lane, ba.workgroup = 2, 256
    ba.run()
    assert np.array_equal(ba.lost, first)
    assert np.sum(first < turns) > 0


def test_autotuner_picks_a_candidate():
    turns = 128
    _, _, _, _, _, ba, _ = _setup(4, (64, 300), turns)
    ba.block_turns = None
    ba.run()
    assert ba.block_turns in (32, 64, 128) and ba.waves_per_simd in (2, 4, 8)
    assert ba.particles_per_lane in (1, 2) and ba.workgroup in (64, 256)
    assert np.all(ba.lost <= turns)


def test_seed_closed_orbits_are_fixed_points():
    """Per-seed Newton closed orbits (tangent-kernel Jacobian, tolerance
    1e-13 m) are fixed points of the oracle's one-turn map with that seed's
    errors bound as per-particle parameters."""
    from ufo_amd.batch_adjoint import SeedClosedOrbit
    ring, prm, vals, counts, fam = lattices.dba_family_model(seeds=16)
    co = SeedClosedOrbit(ring, ufo.KICK | ufo.DOUBLE_PRECISION, prm, vals, tol=1e-13)
    orbits = co.run()
    print("orbits", orbits, "steps", co.steps)
    assert co.converged.all() and co.steps.max() <= 8
    pool = np.concatenate([orbits, vals], axis=1)
    t = runner.track(ring, pool, flags=ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED, turns=1,
                     parameters=["x", "px", "y", "py"] + prm, where=[-1])[:, 0, :4]
    assert np.max(np.abs(t - orbits)) < 1e-13
    assert np.max(np.abs(orbits)) > 1e-7


def test_batched_adjoint_gradients_per_seed():
    """6D tracking of 3 seeds in one launch: final states agree with the
    oracle (per-particle parameters), per-seed gradients of the per-seed
    amplitude loss with respect to all 8 sextupole families (both period
    sets, so the set-B accumulators of ``ufo_gmap`` take part) agree with
    central differences."""
    import torch
    from ufo_amd import adjoint
    from ufo_amd.batch_adjoint import BatchedAdjoint
    ring, prm, vals, counts, fam = lattices.dba_family_model(seeds=3, count_range=(150, 300))
    turns = 96
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    coords = ["x", "px", "y", "py", "dp"]
    rng = np.random.default_rng(9)
    n = int(counts.sum())
    x0 = np.stack([rng.normal(0, 4e-3, n), rng.normal(0, 1e-4, n), rng.normal(0, 2e-3, n),
                   rng.normal(0, 1e-4, n), rng.normal(0, 2e-3, n)], 1)
    seed_of = np.repeat(np.arange(len(counts)), counts)
    ba = BatchedAdjoint(ring, flags, turns, coords, prm, vals, counts, fam)
    assert ba.kbwd.src.count("ufo_gmap[rep]") > 0
    xf, lost = ba.forward(x0)
    xf_np = xf.cpu().numpy()
    ref = runner.track(ring, np.concatenate([x0, vals[seed_of]], axis=1), flags=flags, turns=turns,
                       parameters=coords + prm, where=[-1])[:, -1, :4]
    ok = np.isfinite(ref[:, 0]) & np.isfinite(xf_np[:, 0])
    assert np.array_equal(np.isfinite(ref[:, 0]), np.isfinite(xf_np[:, 0]))
    rel = np.max(np.abs(xf_np[ok] - ref[ok])) / np.max(np.abs(ref[ok]))
    print("forward max rel diff", rel, "alive", ok.sum(), "of", n)
    assert rel < 1e-10

    def per_seed_loss(xf_t):
        return torch.stack([adjoint.amplitude_loss(xf_t[torch.as_tensor(seed_of == s, device=xf_t.device)])
                            for s in range(len(counts))])

    xf_t = xf.clone().requires_grad_(True)
    per_seed_loss(xf_t).sum().backward()
    _, gseed = ba.backward(xf_t.grad)
    g = gseed.cpu().numpy()
    # 4th-order central differences (steps +-h, +-2h) with h = 1e-3 on k2 ~ 38:
    # the truncation error is far below 1e-6 and the rounding noise of 96
    # turns of 6D tracking (~1e-15 in the loss) is divided by a step ten
    # times larger than a 2nd-order difference could use (at h = 1e-4 a
    # family gradient of 4e-10 carried 3e-6 of noise)
    h = 1e-3
    col = [prm.index(f) for f in fam]
    fd = np.zeros_like(g)
    for j, c in enumerate(col):
        # seeds are independent: one set of forwards moves family j of
        # every seed and gives every seed's difference
        for k, w in ((1, 8.0), (2, -1.0)):
            vp, vm = vals.copy(), vals.copy()
            vp[:, c] += k * h
            vm[:, c] -= k * h
            lp = per_seed_loss(BatchedAdjoint(ring, flags, turns, coords, prm, vp, counts, fam).forward(x0)[0])
            lm = per_seed_loss(BatchedAdjoint(ring, flags, turns, coords, prm, vm, counts, fam).forward(x0)[0])
            fd[:, j] += w * (lp - lm).cpu().numpy()
        fd[:, j] /= 12.0 * h
    for j in range(len(col)):
        for s in range(len(counts)):
            print(f"seed {s} family {fam[j]}: adjoint {g[s, j]:.10e} fd {fd[s, j]:.10e}")
            assert abs(g[s, j] - fd[s, j]) <= 1e-6 * max(abs(fd[s, j]), 1e-3 * np.abs(fd[:, j]).max())


def test_batched_config_full_size_sampled():
    """The full error-seed workload: 256 seeds of the 4-period DBA ring,
    per-seed counts uniform in [4096, 32768] (4.9 M particles), 2048 turns
    in 64-turn graph blocks.  A sample of 384 particles (every seed
    represented, a third of them particles the GPU reports lost) is run
    through the oracle with the seed's errors as per-particle parameters:
    loss turns are identical on every particle of the sample whose loss turn
    is regular (strict, fast-math and one-ulp-moved oracle runs agree, and
    their trajectories stay within 1e-9 of each other: over 2048 turns
    particles at the chaotic edge of the aperture survive all three runs by
    chance while their trajectories part by O(1))."""
    turns = 2048
    ring, prm, vals, counts = lattices.dba_error_model(periods=4, seeds=256)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    ba = BatchedAperture(ring, flags, turns, prm, vals, counts, block_turns=64)
    pool = np.concatenate([lattices.seed_grid(int(c)) for c in counts])
    ba.parameters[:] = pool
    ba.run()
    lost = ba.lost.astype(np.int64)
    seed_of = np.repeat(np.arange(len(counts)), counts)
    rng = np.random.default_rng(11)
    gone = np.flatnonzero(lost < turns)
    pick = np.concatenate([rng.choice(len(pool), 256, replace=False), rng.choice(gone, 128, replace=False)])
    full = np.concatenate([pool[pick], vals[seed_of[pick]]], axis=1)
    strict, regular = loss_turn_runs(
        lambda p, m: runner.stable_aperture(ring, p, flags=flags, turns=turns, parameters=["x", "y"] + prm, mode=m),
        full,
        track_fn=lambda p, m: runner.track(ring, p, flags=flags, turns=turns, parameters=["x", "y"] + prm,
                                           where=[-1], mode=m))
    print(f"particles {len(pool)}, GPU lost {len(gone)}; sample {len(pick)}: oracle lost "
          f"{np.sum(strict < turns)}, regular {regular.sum()}, differ {np.sum(lost[pick] != strict)}")
    assert regular.sum() > 0.7 * len(pick)
    assert np.array_equal(lost[pick][regular], strict[regular])


def test_batched_empty_seeds():
    """Seeds with no particles between and around non-empty ones, and a
    batch with no particles at all: the loss turns equal those of the plain
    StableAperture kernel with each seed's errors as per-particle
    parameters, bit for bit."""
    turns = 200
    ring, prm, vals, _ = lattices.dba_error_model(periods=4, seeds=6)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    counts = np.array([0, 150, 0, 0, 77, 0])
    ba = BatchedAperture(ring, flags, turns, prm, vals, counts, block_turns=64)
    pool = np.concatenate([lattices.seed_grid(int(c), xlim=0.03, ymax=0.015) for c in counts if c])
    ba.parameters[:] = pool
    ba.run()
    seed_of = np.repeat(np.arange(len(counts)), counts)
    sa = ufo.StableAperture(ring, flags=flags, turns=turns, particles=len(pool), parameters=["x", "y"] + prm)
    sa.parameters[:] = np.concatenate([pool, vals[seed_of]], axis=1)
    sa.run()
    assert 0 < np.sum(sa.lost < turns) < len(pool)
    assert np.array_equal(ba.lost, sa.lost)
    empty = BatchedAperture(ring, flags, turns, prm, vals[:2], [0, 0], block_turns=64)
    empty.run()
    assert empty.lost.shape == (0,)


def test_batched_adjoint_all_families_full_size():
    """The batched-adjoint workload at full size: 64 seeds x 8192 particles
    of the 4-period DBA ring with errors (sigma_x = 4 mm, sigma_y = 2 mm,
    dp sigma 2e-3, 6D), launched around each seed's Newton closed orbit,
    1024 turns; per-seed gradients with respect to all 8 sextupole families
    against central differences, every (seed, family) pair.

    Over 1024 turns a particle near the dynamic-aperture edge is chaotic:
    its final state moves by O(1) under a change of the strengths far above
    any usable difference step, so no step makes a difference quotient of a
    loss that contains it accurate.  The objective is therefore the
    per-seed amplitude loss over the bunch's regular core, the surviving
    particles whose final state amplifies a change of the initial one by at
    most 30 (|J^T u| / |u| <= 30 for the direction u of dL/dx_final, J the
    particle's 1024-turn Jacobian, taken from the adjoint's own dL/dx0):
    about 277 000 of the 524 000 survivors.  Differences of 4th order
    (steps +-h, +-2h, h = 1e-3 on k2 ~ 38) and of 6th order (h = 3e-3)
    agree with the adjoint to ~5e-8 of each family's largest gradient on
    this set (a sweep over h = 1e-4 .. 3e-3 shows the plateau; with the
    mask at 100 the same differences reach only ~4e-6, at 1000 ~1)."""
    import torch
    from ufo_amd import batch
    from ufo_amd.batch_adjoint import BatchedAdjoint, SeedClosedOrbit
    seeds, per_seed, turns = 64, 8192, 1024
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    coords = ["x", "px", "y", "py", "dp"]
    ring, prm, vals, _, fam = lattices.dba_family_model(seeds=seeds)
    counts = np.full(seeds, per_seed)
    co = SeedClosedOrbit(ring, flags, prm, vals, tol=1e-13)
    orbits = co.run()
    assert co.converged.all()
    rng = np.random.default_rng(4)
    n = seeds * per_seed
    x0 = np.stack([rng.normal(0, 4e-3, n), rng.normal(0, 1e-4, n), rng.normal(0, 2e-3, n),
                   rng.normal(0, 1e-4, n), rng.normal(0, 2e-3, n)], 1)
    seed_of = np.repeat(np.arange(seeds), per_seed)
    x0[:, :4] += orbits[seed_of]
    ba = BatchedAdjoint(ring, flags, turns, coords, prm, vals, counts, fam)
    dev = ba.device
    x0 = torch.tensor(x0, device=dev)
    seg = torch.tensor(seed_of, device=dev)
    xf0, _ = ba.forward(x0)
    alive = torch.isfinite(xf0[:, 0])
    lam = torch.zeros_like(xf0)
    lam[:, 0] = 2 * torch.nan_to_num(xf0[:, 0])
    lam[:, 2] = 2 * torch.nan_to_num(xf0[:, 2])
    gx0, _ = ba.backward(lam)
    amp = gx0.norm(dim=1) / lam.norm(dim=1).clamp_min(1e-300)
    mask = alive & (amp <= 30.0)

    def per_seed_loss(xf):
        r2 = torch.where(mask, xf[:, 0] ** 2 + xf[:, 2] ** 2, torch.zeros_like(xf[:, 0]))
        soft = torch.nn.functional.softplus(20.0 * (r2 / 0.03 ** 2 - 0.25)) / 20.0
        pen = torch.where(mask, soft, torch.zeros_like(r2))
        return torch.zeros(seeds, dtype=torch.float64, device=dev).index_add_(0, seg, r2 + 1e-5 * pen) / per_seed

    xf_t = xf0.clone().requires_grad_(True)
    per_seed_loss(xf_t).sum().backward()
    g = ba.backward(torch.nan_to_num(xf_t.grad, nan=0.0))[1].cpu().numpy()
    probe = BatchedAdjoint(ring, flags, turns, coords, prm, vals, counts, fam)
    probe._buffers()

    def loss_at(v):
        probe._dev["table"] = torch.tensor(batch.build_table(probe.layout, v), dtype=torch.float64, device=dev)
        xf, _ = probe.forward(x0)
        assert bool(torch.isfinite(xf[mask, 0]).all())
        return per_seed_loss(xf).cpu().numpy()

    weights = {4: (1e-3, {1: 8.0, 2: -1.0}, 12.0), 6: (3e-3, {1: 45.0, 2: -9.0, 3: 1.0}, 60.0)}
    worst = {}
    for order, (h, w, den) in weights.items():
        fd = np.zeros_like(g)
        for j, f in enumerate(fam):
            c = prm.index(f)
            for k, wk in w.items():
                vp, vm = vals.copy(), vals.copy()
                vp[:, c] += k * h
                vm[:, c] -= k * h
                fd[:, j] += wk * (loss_at(vp) - loss_at(vm))
            fd[:, j] /= den * h
        crit = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-3 * np.max(np.abs(fd), axis=0))
        worst[order] = crit.max(axis=0)
        print(f"order {order}, h {h:g}: regular core {int(mask.sum())} of {int(alive.sum())} survivors; max "
              f"|adjoint - fd| / |fd| per family " + " ".join(f"{fam[j][0]} {worst[order][j]:.1e}" for j in range(8)))
    for order in weights:
        assert worst[order].max() <= 1e-6, (order, worst[order])
