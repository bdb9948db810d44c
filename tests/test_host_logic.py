"""Host logic of the drivers that needs no GPU: parameter staging, kernel
argument checks, launch-shape validation."""
import numpy as np
import pytest
import torch

import ufo_amd as ufo
from ufo_amd import lattices, runtime
from ufo_amd.sim import Simulation


class _Stub(Simulation):
    def __init__(self, shape):
        self._init_common(ufo.DOUBLE_PRECISION, None, None)
        self._pool_host = torch.zeros(shape, dtype=torch.float64)
        self.parameters = self._pool_host.numpy()


def test_parameters_filled_in_place_are_the_pool():
    s = _Stub((4, 2))
    s.parameters[:, 0] = [1, 2, 3, 4]
    assert s._host_pool().numpy()[:, 0].tolist() == [1, 2, 3, 4]


def test_rebound_parameters_are_copied_into_the_pool():
    """The reference tracks whatever array ``parameters`` refers to at run()
    (ufo/sa.py:141 copies self.parameters); rebinding the attribute works."""
    s = _Stub((4, 2))
    s.parameters = np.arange(8.0).reshape(4, 2)
    assert np.array_equal(s._host_pool().numpy(), np.arange(8.0).reshape(4, 2))
    s.parameters = np.arange(8.0)            # same size, flat
    assert np.array_equal(s._host_pool().numpy(), np.arange(8.0).reshape(4, 2))
    s.parameters = np.zeros((3, 2))
    with pytest.raises(ValueError):
        s._host_pool()


def test_kernel_arguments_are_checked():
    with pytest.raises(ValueError):
        runtime.Module._args([torch.zeros(3)])                   # host tensor
    with pytest.raises(ValueError):
        runtime.Module._args([2 ** 31])                          # does not fit int
    arr, holders = runtime.Module._args([5, 2.5, True])
    assert len(holders) == 3


def test_batched_launch_shape_validation():
    from ufo_amd.batch import batch_kernels, _paired
    ring, prm, vals, counts = lattices.dba_error_model(periods=4, seeds=2)
    k, lay = batch_kernels(ring, ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED, ["x", "y"], prm)
    assert "ufo_bt_turns2" in k.src and "ufo_bt_turns(" in k.src
    # the paired body renames coordinates and temporaries, shares loop structure
    out = _paired(["for (int rep = 0; rep < 4; ++rep) {", "x = fma(px, 0.5, x);", "{ ufloat t1 = x * y; }", "}"])
    assert out == ["for (int rep = 0; rep < 4; ++rep) {", "x = fma(px, 0.5, x);", "xb = fma(pxb, 0.5, xb);",
                   "{ ufloat t1 = x * y; }", "{ ufloat tb1 = xb * yb; }", "}"]


def test_batched_counts_are_checked():
    """One non-negative particle count per seed, at least one seed; seeds
    with no particles are allowed (the device side needs none)."""
    from ufo_amd.batch import BatchedAperture
    ring, prm, vals, counts = lattices.dba_error_model(periods=4, seeds=3)
    fl = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    for bad in ([], [5, -1, 2]):
        with pytest.raises(ValueError):
            BatchedAperture(ring, fl, 10, prm, vals[:len(bad)], bad, block_turns=64)
    ba = BatchedAperture(ring, fl, 10, prm, vals, [0, 7, 0], block_turns=64)
    assert ba.particles == 7 and list(ba.offsets) == [0, 0, 7, 7] and ba.parameters.shape == (7, 2)


def test_fp32_host_evaluation_rounds_every_step():
    """ir.evaluate(fp32=True) reproduces an FP32 kernel's evaluation of the
    same expression: every leaf and operation rounded to float32."""
    from ufo_amd import ir
    e = ir.Expr.of("parameters[0]") * 0.3 / 7.0 + ir.Expr.of("parameters[1]") * ir.Expr.of("parameters[0]")
    p = [1.2345678901234567, -0.987654321]
    f32 = np.float32
    expect = f32(f32(f32(p[0]) * f32(0.3)) / f32(7.0)) + f32(f32(p[1]) * f32(p[0]))
    assert ir.evaluate(e, p, None, True) == float(expect)
    assert ir.evaluate(e, p) == (p[0] * 0.3 / 7.0 + p[1] * p[0])


@pytest.mark.parametrize("where", ["every", "quads"])
def test_optics_host_algebra_matches_the_oracle(where):
    """The host half of Optics.run() (periodic Twiss, dispersion and phase
    advance at every observation point from the three rays' linear map),
    fed the oracle's ALBA linear-map track, against the oracle's restatement
    of ufo/optics.py:156-238 on the same track: tunes and Twiss functions
    agree to rounding."""
    import os
    from oracle import runner
    from conftest import OPTICS_DIR
    ring = ufo.Lattice(path=os.path.join(OPTICS_DIR, "alba.mad")).RING
    flat = ring.flatten()
    idx = list(range(len(flat))) if where == "every" else ring.find(lambda e: isinstance(e, ufo.Quadrupole))
    ref = runner.Optics(ring, where=list(idx), flags=ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION)
    ref.run()
    o = object.__new__(ufo.Optics)
    o.periodic, o.count = True, ref.count
    o.dy0 = o.dpy0 = 0.0    # as Optics.__init__ leaves them (no vertical dispersion source)
    o.ax, o.ay, o.bx, o.by, o.dx, o.dy, o.dpx, o.dpy, o.mux, o.muy = (np.empty(ref.count) for _ in range(10))
    o._twiss(np.asarray(ref.track, dtype=np.float64))
    assert abs(o.qx - ref.qx) < 1e-14 and abs(o.qy - ref.qy) < 1e-14
    for name in ("bx", "by", "ax", "ay", "dx", "dpx", "mux", "muy"):
        a, b = getattr(o, name), np.asarray(getattr(ref, name))
        scale = max(1.0, float(np.max(np.abs(b))))
        assert np.max(np.abs(a - b)) < 1e-12 * scale, name


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and a
    JSON line saying so, before touching a device."""
    import json
    import os
    import subprocess
    import sys
    from conftest import ROOT
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


@pytest.mark.parametrize("n", [65, 100, 128, 129, 4096, 4097, 10000])
def test_claim_order_is_a_permutation_of_whole_runs(n):
    """StableAperture's claim order (computed on the device, here on a host
    tensor): a permutation of the particles made of runs of CLAIM_RUN
    consecutive particles, the full runs by increasing normalised
    amplitude and a short last run last; NaN coordinates are accepted;
    pool order (None) for pools of fewer than two runs and with
    claim_order None."""
    sa = ufo.StableAperture(lattices.fodo_ring(2), flags=ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED, turns=4,
                            particles=n, parameters=["x", "y"])
    rng = np.random.default_rng(n)
    pool = torch.tensor(rng.normal(size=(n, 2)) * np.linspace(0.1, 3.0, n)[:, None])
    pool[7, 0] = float("nan")
    sa._pool_dev = pool
    if n // sa.CLAIM_RUN < 2:
        assert sa.order_device() is None
        return
    order = sa.order_device().numpy()
    assert order.dtype == np.int32 and np.array_equal(np.sort(order), np.arange(n))
    R = sa.CLAIM_RUN
    starts = [i for i in range(len(order)) if order[i] % R == 0]
    runs = [order[a:b] for a, b in zip(starts, starts[1:] + [len(order)])]
    assert all(np.array_equal(r, np.arange(r[0], r[0] + len(r))) for r in runs)
    full = n // R * R
    v = np.nan_to_num(pool.numpy()[:full])
    key = ((v / np.abs(v).max(axis=0)) ** 2).sum(axis=1)
    sums = [key[r].sum() for r in runs if r[0] < full]
    assert np.all(np.diff(sums) >= -1e-9)
    if full < n:
        assert np.array_equal(order[full:], np.arange(full, n))
    sa.claim_order = None
    assert sa.order_device() is None
    small = ufo.StableAperture(lattices.fodo_ring(2), flags=ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED, turns=4,
                               particles=64, parameters=["x", "y"])
    small._pool_dev = torch.zeros((64, 2), dtype=torch.float64)
    assert small.order_device() is None
