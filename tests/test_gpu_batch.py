"""GPU checks of the batched (multi-seed) dynamic-aperture tracker against
the oracle, where each particle carries its seed's error values as
per-particle parameters (the reference's way of batching lattice variants,
HOWTO.md:20-22)."""
import numpy as np
import pytest

import ufo_amd as ufo
from ufo_amd import lattices
from ufo_amd.batch import BatchedAperture
from oracle import runner

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
    strict = runner.stable_aperture(ring, full, flags=flags, turns=turns, parameters=["x", "y"] + prm)
    fast = runner.stable_aperture(ring, full, flags=flags, turns=turns, parameters=["x", "y"] + prm, mode="fast")
    regular = strict == fast
    lost = ba.lost.astype(np.int64)
    print(f"particles {len(pool)}, lost {np.sum(strict < turns)}, regular {regular.sum()}, "
          f"differ {np.sum(lost != strict)}")
    assert 0.2 * len(pool) < np.sum(strict < turns) < 0.9 * len(pool)
    assert np.array_equal(lost[regular], strict[regular])


def test_graph_replay_and_block_size_do_not_change_results():
    turns = 256
    _, _, _, counts, _, ba, pool = _setup(5, (64, 700), turns)
    ba.run()
    first = ba.lost.copy()
    ba.run()                                # replays the captured graph
    assert np.array_equal(ba.lost, first)
    ba.block_turns = 32
    ba.run()                                # another block size: same physics
    assert np.array_equal(ba.lost, first)
    assert np.sum(first < turns) > 0


def test_autotuner_picks_a_candidate():
    turns = 128
    _, _, _, _, _, ba, _ = _setup(4, (64, 300), turns)
    ba.block_turns = None
    ba.run()
    assert ba.block_turns in (32, 64, 128) and ba.waves_per_simd in (4, 8)
    assert np.all(ba.lost <= turns)
