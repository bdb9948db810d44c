"""ORACLE — test infrastructure only.

A CPU restatement of the reference's kernels (``refgen``: the text generator,
``runner``: gcc build + ctypes calls).  Imported by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
only, and there only as the checker; the product never imports it.

How each part is pinned (no part is left unpinned):
  runner.track          MAD-X tracking vector (tests/tracking.py), mpmath
                        closed forms of the kick series, exact drift and
                        wire, N-turn linear FODO = matrix power, thick bend
                        vs RK4, y-mirror symmetry, time reversal, the
                        octupole ring of tests/oct.py (200-bit mpmath)
  runner.stable_aperture  the HOWTO 12 x 12 ALBA grid (HOWTO.md:332-370)
  runner.closed_orbit   the HOWTO FP64 closed-orbit table (HOWTO.md:494-503),
                        to every printed digit
  runner.Optics         ALBA tunes and finite-difference chromaticity of
                        tests/optics.py and tests/optics3.py (MAD-X),
                        thick-vs-teapot tunes, periodic = propagated
                        (tests/optics2.py), misaligned quad = bend
                        (tests/align.py), sampled Twiss integrals
(tests/test_oracle_pins.py, tests/test_oracle_math.py,
tests/test_reference_cases.py, tests/test_optics_consumers.py)
"""
