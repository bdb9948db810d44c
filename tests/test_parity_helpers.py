"""The parity measure of tests/parity.py on constructed tracks (CPU)."""
import numpy as np

from parity import particle_err


def _tracks(n=8, samples=5, seed=1):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 1e-3, (n, samples, 6))


def test_diverging_particle_does_not_hide_other_errors():
    """A particle that has grown to 1e200 (finite, so not yet lost) must not
    raise the error floor of the others: a 1e-9 relative change of another
    particle stays visible."""
    b = _tracks()
    b[3] = 1e200
    a = b.copy()
    a[0] *= 1.0 + 1e-9
    err = particle_err(a, b)
    assert 0.5e-9 < err[0] < 2e-9
    assert err[3] == 0.0 and np.all(err[1:3] == 0.0)


def test_zero_crossings_are_measured_against_the_coordinate_rms():
    """An entry of the oracle at exactly zero is compared against the rms of
    its coordinate, not against zero."""
    b = _tracks()
    b[2, 1, 0] = 0.0
    a = b.copy()
    a[2, 1, 0] = 1e-18
    rms = np.sqrt(np.mean(b[:, :, 0] ** 2))
    np.testing.assert_allclose(particle_err(a, b)[2], 1e-18 / rms, rtol=1e-12)


def test_lost_entries_are_excluded():
    """Entries where the oracle is NaN (lost) do not count."""
    b = _tracks()
    b[5, 2:] = np.nan
    a = b.copy()
    a[5, 2:] = 123.0
    assert particle_err(a, b)[5] == 0.0
