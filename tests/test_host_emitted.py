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
