"""Host-side checks of the lowering and of the HIP emitter (CPU only):
flop accounting, period detection, and that every generated kernel compiles
with hipRTC for gfx950 (no device needed)."""
import os

import numpy as np
import pytest

import ufo_amd as ufo
from ufo_amd import codegen, ir, lattices, runtime

from conftest import OPTICS_DIR
from test_gpu_parity import _all_elements_line


def test_fodo_flops_and_period():
    ring = lattices.fodo_ring(16)
    k = codegen.sa_kernel(ring, ufo.KICK | ufo.DOUBLE_PRECISION, 1000, ["x", "px", "y", "py", "dp"], 0.0)
    assert k.flops_per_turn == 1600          # cell: 2 x (5 drifts x 4 + 4 kicks x 4) + 2 x 10 + 2 x 4
    assert k.elements == 96
    head, p, reps = k.period
    assert head == 0 and reps == 16


def test_dba_flops():
    ring = lattices.dba_ring(16)
    k = codegen.sa_kernel(ring, ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED, 1024, ["x", "y"], 0.0)
    # per period: 2 bends x 84 + 6 quads x 68 + 4 sextupoles x 60 + 13 drifts x 4
    assert k.flops_per_turn == 16 * (2 * 84 + 6 * 68 + 4 * 60 + 13 * 4)
    assert k.elements == 400
    assert k.period[2] == 16


def test_kick_flop_rule():
    E = ir.E
    assert ir.kick_flops([E(0.0), E(0.3)], []) == 4               # quad slice
    assert ir.kick_flops([E(0.0), E(0.0), E(1.0)], []) == 10      # sextupole slice
    assert ir.kick_flops([E(1.0)], [E(2.0)]) == 2                 # dipole both planes
    assert ir.kick_flops([E(0), E(1), E(1)], [E(0), E(0), E(1)]) == 6 + 6 + 1 + 2


def test_fold_identities():
    p = ir.Expr.of("parameters[2]")
    assert p.t == ("p", 2)
    assert ir.fold(p * 0.0).is_const() and ir.fold(p * 0.0).value == 0.0
    assert ir.fold(p * 1.0) == p
    assert ir.fold(ir.E(0.5) * 4.0 / 8.0).value == 0.25
    assert ir.fold(ir.OODPPO * 2.0, env={"oodppo": 0.5}).value == 1.0
    assert ir.fold(ir.select(ir.gt(ir.E(1.0), 0.0), 3.0, 4.0)).value == 3.0


def test_find_period_with_head_and_tail():
    a, b, c = ir.Drift(ir.E(1.0)), ir.Drift(ir.E(2.0)), ir.Weak(ir.E(0.1))
    ops = [c] + [a, b] * 5 + [c, c]
    assert codegen.find_period(ops) == (1, 2, 5)


def test_line_method_fracture_and_binding():
    """ufo/lattice.py:86-119: fracture splits, parameter binding by label
    and by flat index."""
    q = ufo.Quadrupole("Q", length=1.0, k1=1.0, slices=2)
    d = ufo.Drift("D", length=2.0)
    line = ufo.Line("L", [q, d, q])
    pieces = line.method(fracture=[0.5, 1, 2.25], flags=ufo.KICK, parameters=[("Q", "k1"), (1, "length")])
    assert len(pieces) == 4
    kicks = [op for piece in pieces for op in piece if isinstance(op, ir.Kick)]
    assert all(("p", 0) in [leaf.t for leaf in k.kn[1].leaves()] for k in kicks)
    drifts = [op for op in pieces[1] + pieces[2] if isinstance(op, ir.Drift)]
    assert any(("p", 1) in [leaf.t for leaf in dr.L.leaves()] for dr in drifts)


@pytest.mark.parametrize("flags", [ufo.KICK | ufo.DOUBLE_PRECISION, ufo.DOUBLE_PRECISION | ufo.EXACT,
                                   ufo.KICK | ufo.FIVED, ufo.LINEAR | ufo.ACHROMATIC])
def test_generated_kernels_compile(flags):
    line = _all_elements_line()
    for k in (codegen.sa_kernel(line, flags, 10, ["x", "px", "y", "py", "dp"], 1e-3),
              codegen.track_kernel(line, flags, 10, ["x", "px", "y", "py", "dp"], [0.5, 2, -1], 1e-3),
              codegen.optics_kernel(line, flags, [0.5, 4, -1], ["dp"])):
        code = runtime.compile_only(k.src, codegen.HIPRTC_OPTIONS, "gfx950")
        assert len(code) > 1000


def test_alba_kernels_compile():
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    k = codegen.sa_kernel(alba.RING, ufo.FIVED | ufo.DOUBLE_PRECISION, 100, ["x", "y"], 0.0)
    assert k.period[2] == 4     # 4 identical cells
    runtime.compile_only(k.src, codegen.HIPRTC_OPTIONS, "gfx950")


def test_compile_error_is_reported():
    with pytest.raises(runtime.UfoError, match="hipRTC"):
        runtime.compile_only("extern \"C\" __global__ void k() { this is not C++; }", "", "gfx950")


def test_aperture_window_spliced_verbatim():
    line = ufo.Line("l", [ufo.Aperture("ap", window="x > 0.1 && y > 0.1")])
    k = codegen.track_kernel(line, ufo.DOUBLE_PRECISION, 1, ["x", "y"], [-1], 0.0)
    assert "if (x > 0.1 && y > 0.1 == 0) x = UFO_NAN;" in k.src


def test_strided_period_for_per_magnet_parameters():
    """Per-sextupole parameters (the HOWTO closed-orbit example) keep the
    4-cell loop of ALBA: repetitions differ by a parameter stride of 30."""
    from ufo_amd import tangent
    alba = ufo.Lattice(os.path.join(OPTICS_DIR, "alba.mad"))
    sext = alba.RING.find(lambda e: isinstance(e, ufo.Sextupole))
    ops = [op for piece in alba.RING.method(flags=ufo.FIVED, parameters=[(e, "dx") for e in sext])
           for op in piece]
    head, p, reps, stride = codegen.find_period(ops, with_stride=True)
    assert reps == 4 and stride == 30
    k = tangent.co_kernel(alba.RING, ufo.FIVED | ufo.DOUBLE_PRECISION, [(e, "dx") for e in sext], 64, 1e-15)
    assert "row[" in k.src and "rep * 30]" in k.src
    runtime.compile_only(k.src, codegen.HIPRTC_OPTIONS, "gfx950")


def test_tangent_kernels_compile():
    from ufo_amd import tangent
    line = _all_elements_line()
    for fl in (ufo.KICK | ufo.DOUBLE_PRECISION, ufo.DOUBLE_PRECISION | ufo.EXACT, ufo.DOUBLE_PRECISION | ufo.FIVED):
        runtime.compile_only(tangent.jvp_kernel(line, fl, 2, ["x", "px", "y", "py", "dp"], 1e-3).src,
                             codegen.HIPRTC_OPTIONS, "gfx950")
        runtime.compile_only(tangent.co_kernel(line, fl, ["x"], 20, 1e-15).src, codegen.HIPRTC_OPTIONS, "gfx950")


def test_z_function_matches_brute_force():
    """The linear-time longest-common-prefix table behind the period search
    equals the direct definition on random sequences over small alphabets
    (long repeats) and on periodic ones."""
    from ufo_amd.codegen import _z_function
    rng = np.random.default_rng(5)
    seqs = [list(rng.integers(0, k, n)) for k in (1, 2, 3, 7) for n in (1, 2, 5, 40, 257)]
    seqs += [[0, 1, 2] * 30 + [0, 1], [4, 4, 5] * 17 + [9]]
    for s in seqs:
        z = _z_function(s)
        for i in range(1, len(s)):
            k = 0
            while i + k < len(s) and s[k] == s[i + k]:
                k += 1
            assert z[i] == k, (s, i)


def test_gradient_accumulators_cleared_behind_a_barrier():
    """The reverse kernels clear their LDS gradient accumulators and then
    wait at a workgroup barrier before any wave adds to them (without it a
    wave that starts late can clear a slot another wave has already added
    to), and wait again before the per-workgroup sums are read."""
    from ufo_amd import adjoint, batch_adjoint
    ring = lattices.fodo_ring(4, distinct=True)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    coords = ["x", "px", "y", "py", "dp"]
    shared = [("QF0", "k1"), ("SD1", "k2")]
    kb = adjoint.adjoint_kernels(ring, flags, 20, coords, shared, every=8)[1]
    kbb = batch_adjoint.batched_adjoint_kernels(ring, flags, 20, coords, shared, shared, every=8)[1]
    for src in (kb.src, kbb.src):
        clear = src.index("acc[i] = 0.0")
        first_add = src.index("unsafeAtomicAdd(&acc[")
        last_add = src.rindex("unsafeAtomicAdd(&acc[")
        read = src.index("s += acc[")
        assert "__syncthreads()" in src[clear:first_add]
        assert "__syncthreads()" in src[last_add:read]
