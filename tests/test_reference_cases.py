"""CPU counterparts of the reference's remaining test scripts: every element
of tests/elements.py's lattice through ``method()`` and the emitted kernel,
the octupole ring of tests/oct.py against an mpmath closed form, and the
periodic/propagated Twiss consistency of tests/optics2.py."""
import os

import mpmath
import numpy as np
import pytest

import ufo_amd as ufo
from ufo_amd import codegen
from oracle import runner

from conftest import OPTICS_DIR
import host_emulation
from parity import assert_parity, compare_tracks, reference_runs, summary

COORD6 = ["x", "px", "y", "py", "dp"]


def lattice(name):
    return ufo.Lattice(os.path.join(OPTICS_DIR, f"{name}.mad"))


@pytest.mark.parametrize("flags", [0, ufo.DOUBLE_PRECISION, ufo.KICK | ufo.DOUBLE_PRECISION,
                                   ufo.KICK | ufo.DOUBLE_PRECISION | ufo.EXACT,
                                   ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION])
def test_elements_lattice_method_and_emitted_kernel(flags):
    """tests/elements.py loads elements.mad (quadrupole, sbend with pole
    faces, drift, marker, sextupole, octupole multipole) and calls
    TEST.method(); here the method pieces are also emitted as a Track kernel,
    built for the host and compared with the oracle over 5 passes."""
    lat = lattice("elements")
    pieces = lat.TEST.method(flags=flags)
    assert len(pieces) == 1 and len(pieces[0]) > 0
    rng = np.random.default_rng(11)
    pool = np.stack([rng.normal(0, 1e-3, 64), rng.normal(0, 1e-4, 64), rng.normal(0, 1e-3, 64),
                     rng.normal(0, 1e-4, 64), rng.normal(0, 1e-3, 64)], 1)
    line = lat.TEST
    # a fractional position on the marker (element 3) is never written: the
    # reference's tracks hold len(where) * turns samples, the first
    # dumps-per-turn * turns of them written in dump order, the rest undefined
    k = codegen.track_kernel(line, flags, 5, COORD6, [1, 3.5, -1], 0.0)
    assert k.dumps_per_turn == 2
    host = host_emulation.run_track(k, pool, 5 * 3)

    def oracle(p, mode):
        return runner.track(line, p, flags=flags, turns=5, parameters=COORD6, where=[1, 3.5, -1], mode=mode)

    strict, fast, moved = reference_runs(oracle, pool, fp32=not (flags & ufo.DOUBLE_PRECISION))
    r = compare_tracks(host, strict, fast, moved)
    print(f"flags {flags}: " + summary(r))
    if flags & ufo.DOUBLE_PRECISION:
        assert_parity(r)
    else:
        # FP32, the reference's default precision: no particle is regular at
        # the 3e-13 level, so the bar is the statistical one (as close to
        # the reference as its own fast-math build and one-ulp moves are)
        assert r["flags_ok"]
        assert r["median_gpu"] <= 2.0 * r["median_ref"] and r["p90_gpu"] <= 2.0 * r["p90_ref"]


def _oct_closed_form(x0, turns, k3, L, fp):
    """OCT multipole (KNL[3] = k3) then drift L, y = 0, FIVED (dp = 0):
    px -= k3 x^3 / 3!, x += px L, in mpmath at ``fp`` bits."""
    with mpmath.workprec(fp):
        x, px = mpmath.mpf(x0), mpmath.mpf(0)
        out = []
        for _ in range(turns):
            px = px - mpmath.mpf(k3) * x ** 3 / 6
            x = x + px * mpmath.mpf(L)
            out.append((float(x), float(px)))
    return np.array(out)


@pytest.mark.parametrize("double", [False, True])
def test_octupole_ring_closed_form(double):
    """tests/oct.py: oct.mad (OCT MULTIPOLE KNL={0,0,0,10765.5878}, O DRIFT
    L=0.2876), FIVED, x0 = 0.01, 10 turns, where=[-1] — the oracle and the
    host build of the emitted kernel against the exact map (mpmath, 200
    bits)."""
    lat = lattice("oct")
    flags = ufo.FIVED | (ufo.DOUBLE_PRECISION if double else 0)
    exact = _oct_closed_form(0.01, 10, 10765.5878, 0.2876, 200)
    pool = np.array([[0.01], [0.01]])
    ref = runner.track(lat.RING, pool, flags=flags, turns=10, parameters=["x"], where=[-1])
    k = codegen.track_kernel(lat.RING, flags, 10, ["x"], [-1], 0.0)
    host = host_emulation.run_track(k, pool, 10)
    tol = 1e-13 if double else 2e-5
    for t in (ref, host):
        np.testing.assert_allclose(t[0, :, 0], exact[:, 0], rtol=tol, atol=tol * 0.01)
        np.testing.assert_allclose(t[0, :, 1], exact[:, 1], rtol=tol, atol=tol * 1e-4)


def test_optics_periodic_equals_propagated():
    """tests/optics2.py: the propagated solution started from the periodic
    Twiss values at the end of the ring returns them (ALBA, within 1e-5 as in
    the reference test; FP64 here agrees to 1e-9)."""
    alba = lattice("alba")
    fl = ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION
    per = runner.Optics(alba.RING, where=[-1], periodic=True, flags=fl)
    per.run()
    pro = runner.Optics(alba.RING, where=[-1], periodic=False, flags=fl)
    for a in ("ax", "ay", "bx", "by", "dx", "dpx"):
        setattr(pro, a + "0", getattr(per, a)[-1])
    pro.run()
    for a in ("ax", "ay", "bx", "by", "dx", "dpx"):
        assert abs(getattr(per, a)[-1] - getattr(pro, a)[-1]) < 1e-9 * max(1.0, abs(getattr(per, a)[-1]))


class _NLK(ufo.Multipole):
    """The HOWTO's non-linear injection kicker (HOWTO.md:86-98): a thin
    quadrupole kick switched on at turn 20 only, written as a subclass whose
    method() passes a C-text override that reads the turn counter."""

    def method(self, k1=None, **kwargs):
        k1 = k1 or self.knl[1]
        k1 = f"(turn == 20) ? {k1} : 0."
        return super().method(k1=f"({k1})", **kwargs)


def nlk_ring():
    alba = lattice("alba")
    nlk = _NLK("NLK", knl=[0., 5e-2])
    nlk.oracle_overrides = {"k1": f"((turn == 20) ? {nlk.knl[1]} : 0.)"}   # what method() passes
    alba.RING.insert(0, nlk)
    return alba.RING


def nlk_bunch(count=100, betax=11.26, ex0=4.3e-9):
    rng = np.random.default_rng(0)
    return np.stack([rng.standard_normal(count) * np.sqrt(0.5 * betax * ex0),
                     rng.standard_normal(count) * np.sqrt(0.5 * ex0 / betax)], 1)


def test_howto_time_dependent_kicker():
    """HOWTO.md:86-120 on the host build of the emitted kernel: the kick
    reads the turn counter, the emittance jumps after turn 20 (4.0 -> 4.6
    nm), and the tracks agree with the oracle (FP64, thick ALBA)."""
    ring = nlk_ring()
    pool = nlk_bunch()
    flags = ufo.FIVED | ufo.DOUBLE_PRECISION
    k = codegen.track_kernel(ring, flags, 65, ["x", "px"], [-1], 0.0)
    assert "turn == 20" in k.src
    host = host_emulation.run_track(k, pool, 65)

    def oracle(p, mode):
        return runner.track(ring, p, flags=flags, turns=65, parameters=["x", "px"], where=[-1], mode=mode)

    strict, fast, moved = reference_runs(oracle, pool)
    r = compare_tracks(host, strict, fast, moved)
    print(summary(r))
    # 65 turns of thick ALBA: the reference's own noise exceeds 3e-13 on most
    # particles, so few are regular; the statistical bar covers the rest
    assert_parity(r, min_regular_frac=0.0)
    ex = np.mean(host[:, :, 0] ** 2 / 11.26 + host[:, :, 1] ** 2 * 11.26, axis=0)
    assert abs(ex[19] - ex[0]) < 0.15 * ex[0] and ex[25] > 1.1 * ex[19]
