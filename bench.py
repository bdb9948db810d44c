"""Benchmark: FP64 dynamic-aperture scan on one MI355X.

Workload (one "step"): StableAperture over the 16-period DBA ring (400
elements, teapot KICK, FIVED, dp = 0) for a 1024 x 1024 grid of initial
(x, y), x in [-20, 20] mm, y in [0, 8] mm, 1024 turns, aperture |x|,|y| <
30 mm.  The particle pool is resident in HBM before the timed region.

Metric: particle-element-turns per second, PET = sum over particles of the
turns executed (loss turn + 1, or all turns) x E, E = 400 elements (markers
and apertures excluded).

Also reported:
  roofline   the tracking kernel's canonical FP64 work per launch (executed
             particle-turns x 13888 FLOP, the UFO FLOP rule) over its average
             launch time from HIP events on the launch stream, against the
             MI355X FP64 peak of 78.6 TFLOP/s;
  cpu_baseline  the oracle (the reference's kernels restated in C, gcc
             -O3 -march=native fast-math and strict-IEEE builds, OpenMP over
             the host threads) on an evenly strided sample of the same
             grid, as PET/s, with the host's CPU model and core counts; and
             beside it the reference's own OpenCL kernel on the same GPU
             over the whole grid (``same_gpu_reference_opencl``).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--turns", type=int, default=1024)
    ap.add_argument("--sext-scale", type=float, default=1.0,
                    help="scale of the ring's sextupole strengths for the timed scan (4: the lossy scan)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-seed-adjoint", action="store_true")
    ap.add_argument("--no-fodo", action="store_true")
    ap.add_argument("--track-particles", type=int, default=4096,
                    help="particles of the FODO-16 Track line (4096: the reference's CPU case)")
    ap.add_argument("--no-lossy", action="store_true")
    ap.add_argument("--no-co-optics", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="MiB copied by one device-to-device copy before the timed region "
                         "(reference kernel for HBM traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch
    import ufo_amd as ufo
    from ufo_amd import lattices

    ring = lattices.dba_ring(16, sext_scale=args.sext_scale)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    pool = lattices.da_grid(args.nx, args.ny)
    n = len(pool)
    t0 = time.perf_counter()
    sa = ufo.StableAperture(ring, flags=flags, turns=args.turns, particles=n, parameters=["x", "y"])
    sa.parameters[:] = pool
    sa.module()
    compile_s = time.perf_counter() - t0
    sa.pool_device.copy_(sa._pool_host)
    torch.cuda.synchronize()
    E = sa.kernel.elements
    flop_pt = sa.kernel.flops_per_turn

    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy << 17, dtype=torch.float64, device=sa.device)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        sa.launch()
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ev[2 * k].record(stream)
        sa.launch()
        ev[2 * k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    kernel_ms = [ev[2 * k].elapsed_time(ev[2 * k + 1]) for k in range(args.steps)]

    lost = sa.lost_device.cpu().numpy().view(np.uint32)
    executed = np.minimum(lost.astype(np.int64) + 1, args.turns)
    pt = int(executed.sum())
    pet_step = pt * E
    ms_step = wall / args.steps * 1e3
    value = pet_step * args.steps / wall
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    achieved_tflops = pt * flop_pt / avg_kernel_s / 1e12

    info = sa.module().info(sa.kernel.name)
    out = {
        "metric": "FP64 particle-element-turns/s (DA scan, StableAperture)",
        "value": value,
        "unit": "particle-element-turns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "dynamic-aperture scan: 16-period DBA ring (400 elements, teapot KICK, FIVED"
                        + (f", sextupoles x{args.sext_scale:g}" if args.sext_scale != 1.0 else "") + "), "
                        f"{args.nx}x{args.ny} grid x in [-20,20] mm, y in [0,8] mm, {args.turns} turns",
            "particles": n, "turns": args.turns, "elements": E,
            "flop_per_particle_turn": flop_pt,
            "executed_particle_turns": pt,
            "survivors": int(np.sum(lost == args.turns)),
            "hiprtc_compile_s": round(compile_s, 3),
            "lanes": sa.lanes, "vgprs": info["vgprs"],
            "device": repr(sa.ctx.device),
        },
        "roofline": {
            "bound": "valu-fp64",
            "achieved": achieved_tflops,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP64_TFLOPS,
            "traffic": measured_traffic(args),
            "traffic_source": ("profiles/ufo_sa_dba_da_scan_1M_1024turns.md: rocprofv3 --pmc FETCH_SIZE and "
                               "WRITE_SIZE of this workload, scaled by the counters of a 512 MiB device copy "
                               "(committed profile, not measured in this run)"),
            "kernel": sa.kernel.name,
            "kernel_ms": float(np.mean(kernel_ms)),
            "peak_from_device": sa.ctx.device.fp64_peak / 1e12,
        },
        "cpu_baseline": None,
    }
    _progress(f"DA scan: {float(np.mean(kernel_ms)):.1f} ms per launch")
    if not args.no_lossy:
        out["lossy_da"] = lossy_scan_timing(args.nx, args.ny, args.turns, args.steps, achieved_tflops)
        _progress("lossy DA scan done")
    if not args.no_fodo:
        out["fodo_track"] = fodo_track_timing(particles=args.track_particles,
                                              cpu=not args.no_cpu and args.track_particles == 4096)
        _progress("FODO Track done")
    if not args.no_co_optics:
        out["co_optics"] = co_optics_timing(cpu=not args.no_cpu)
        _progress("ClosedOrbit and Optics done")
    side_cpu = 0.0 if args.no_cpu else args.cpu_seconds / 4
    if not args.no_adjoint:
        out["adjoint"] = adjoint_timing(cpu_seconds=side_cpu)
        _progress("adjoint done")
    if not args.no_batched:
        out["batched"] = batched_timing(cpu_seconds=side_cpu)
        _progress("batched seeds done")
    if not args.no_seed_adjoint:
        out["seed_adjoint"] = seed_adjoint_timing(cpu_seconds=side_cpu)
        _progress("batched adjoint done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ring, flags, pool, args.turns, E, args.cpu_seconds, gpu_lost=lost)
        _progress("CPU baseline done")
    print(json.dumps(out))


def lossy_scan_timing(nx, ny, turns, reps, lossless_tflops, sext_scale=4.0):
    """The DA scan with particle loss: the same ring, grid and turns with
    the sextupole strengths scaled by 4 (k2 = +-152 m^-3), so that the
    dynamic aperture falls inside the grid and about 40 % of the particles
    are lost, most within 50 turns.  Lost lanes refill from the work
    counter; the rate is the canonical FP64 work of the executed
    particle-turns.  ``utilisation_vs_lossless`` is the achieved FP64 rate
    over that of the lossless scan (whose issued and canonical FLOPs are
    equal): the fraction of the FP64 issue the lossy launch turns into
    useful particle-turns (loss, refill and the tail of the launch).  Lanes
    claim particles in StableAperture's default order (runs of 64 by
    increasing amplitude, the sort included in every timed launch);
    ``kernel_ms_pool_order`` times the same launches claiming in pool
    order."""
    import numpy as np
    import torch
    import ufo_amd as ufo
    from ufo_amd import lattices
    ring = lattices.dba_ring(16, sext_scale=sext_scale)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    pool = lattices.da_grid(nx, ny)
    sa = ufo.StableAperture(ring, flags=flags, turns=turns, particles=len(pool), parameters=["x", "y"])
    sa.parameters[:] = pool
    sa.module()
    sa.pool_device.copy_(sa._pool_host)
    sa.launch()
    stream = torch.cuda.current_stream()

    def timed():
        ms = []
        for _ in range(max(1, reps)):
            e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            e[0].record(stream)
            sa.launch()
            e[1].record(stream)
            torch.cuda.synchronize()
            ms.append(e[0].elapsed_time(e[1]))
        return ms

    # the same launches with lanes claiming particles in pool order, for
    # comparison with the default claim order
    order = sa.claim_order
    sa.claim_order = None
    ms_pool = timed()
    sa.claim_order = order
    ms = timed()
    lost = sa.lost_device.cpu().numpy().view(np.uint32)
    executed = np.minimum(lost.astype(np.int64) + 1, turns)
    pt = int(executed.sum())
    k_s = float(np.mean(ms)) / 1e3
    tflops = pt * sa.kernel.flops_per_turn / k_s / 1e12
    return {"workload": f"DA scan with loss: 16-period DBA ring, sextupoles x{sext_scale:g}, {nx}x{ny} grid "
                        f"x in [-20,20] mm, y in [0,8] mm, {turns} turns",
            "kernel_ms": round(float(np.mean(ms)), 3), "pet_per_s": pt * sa.kernel.elements / k_s,
            "lost_fraction": float(np.mean(lost < turns)),
            "executed_particle_turns": pt, "executed_fraction": pt / (len(pool) * turns),
            "loss_turn_quantiles_10_50_90": [float(q) for q in np.percentile(lost[lost < turns], [10, 50, 90])],
            "tflops": tflops, "frac_fp64_peak": tflops / PEAK_FP64_TFLOPS,
            "utilisation_vs_lossless": tflops / lossless_tflops, "lanes": sa.lanes,
            "claim_order": sa.claim_order, "kernel_ms_pool_order": round(float(np.mean(ms_pool)), 3)}


def fodo_track_timing(particles=4096, turns=1000, reps=3, cpu=True):
    """The reference's CPU-runnable case: Track through the 16-cell FODO ring
    (96 elements, KICK, 6D), 4096 particles (x, y ~ N(0, 1 mm), px, py ~
    N(0, 0.1 mrad), dp ~ N(0, 1e-3), seed 0), where = [-1], 1000 turns.
    GPU: kernel time (HIP events) and Track.run() wall time including the
    196.6 MB copy of tracks back to the host, as run() does it (turns split
    over launches, finished samples copied meanwhile) and in one launch
    followed by one copy; CPU: the oracle (gcc fast-math, OpenMP) on the
    whole workload."""
    import numpy as np
    import torch
    import ufo_amd as ufo
    from ufo_amd import lattices
    ring = lattices.fodo_ring(16)
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    coords = ["x", "px", "y", "py", "dp"]
    bunch = lattices.gaussian_bunch(particles, seed=0)
    tr = ufo.Track(ring, flags=flags, turns=turns, particles=particles, parameters=coords, where=[-1])
    tr.parameters[:] = bunch
    tr.run()
    stream = torch.cuda.current_stream()
    k_ms, run_ms, one_ms = [], [], []
    chunks = tr._chunks()
    for _ in range(reps):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        e[0].record(stream)
        tr.launch()
        e[1].record(stream)
        torch.cuda.synchronize()
        k_ms.append(e[0].elapsed_time(e[1]))
    # each way of running, repeated as a caller would: run() as it is, then
    # one launch followed by one copy
    for mode, times in ((None, run_ms), (1, one_ms)):
        tr.copy_chunks = mode
        tr.run()
        for _ in range(reps):
            t = time.perf_counter()
            tr.run()
            times.append(1e3 * (time.perf_counter() - t))
    tr.copy_chunks = None
    alive = np.isfinite(tr.tracks[:, -1, 0])
    nan_turn = np.argmax(~np.isfinite(tr.tracks[:, :, 0]), axis=1)
    pt = int(np.where(alive, turns, nan_turn + 1).sum())
    E = tr.kernel.elements
    out = {"workload": f"16-cell FODO ({E} elements, KICK, 6D), {particles} particles, {turns} turns, where=[-1]",
           "gpu_kernel_ms": round(float(np.median(k_ms)), 3), "gpu_run_ms": round(float(np.median(run_ms)), 3),
           "run_launches": chunks, "gpu_run_one_launch_ms": round(float(np.median(one_ms)), 3),
           "gpu_pet_per_s": pt * E / (float(np.median(k_ms)) * 1e-3), "survivors": int(alive.sum()),
           "gpu_tflops": pt * tr.kernel.flops_per_turn / (float(np.median(k_ms)) * 1e-3) / 1e12}
    out["frac_fp64_peak"] = out["gpu_tflops"] / PEAK_FP64_TFLOPS
    if cpu:
        from oracle import runner
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        runner.track(ring, bunch[:64], flags=flags, turns=turns, parameters=coords, where=[-1], mode="fast")
        t = time.perf_counter()
        runner.track(ring, bunch, flags=flags, turns=turns, parameters=coords, where=[-1], mode="fast")
        dt = time.perf_counter() - t
        out["cpu_ms"] = round(1e3 * dt, 1)
        out["cpu_pet_per_s"] = pt * E / dt
        out["cpu_threads"] = threads
        out["cpu_kind"] = "port (oracle, whole workload)"
    return out


def co_optics_timing(reps=5, cpu=True):
    """ClosedOrbit and Optics on the ALBA ring (768 elements, the lattice of
    the reference's tests and HOWTO): 120 closed orbits, one per sextupole
    displaced by dx = 1 mm (FIVED FP64; Newton on the tangent kernel here,
    the reference's Nelder-Mead search, 200 iterations, in the oracle), and
    the periodic Optics with an observation point at every element (LINEAR |
    ACHROMATIC, FP64).  Each run() is one small launch plus copies and host
    algebra, bound by launch latency rather than by the FP64 units; the GPU
    figures are run() wall times, medians of ``reps`` after one warm-up."""
    import numpy as np
    import ufo_amd as ufo
    lat = ufo.Lattice(path=os.path.join(ROOT, "tests", "golden", "lattices", "alba.mad"))
    ring = lat.RING
    flat = ring.flatten()
    sext = ring.find(lambda e: isinstance(e, ufo.Sextupole))
    prm = [(e, "dx") for e in sext]
    co_flags = ufo.FIVED | ufo.DOUBLE_PRECISION
    co = ufo.ClosedOrbit(ring, particles=len(sext), flags=co_flags, parameters=prm)
    co.parameters[:] = np.identity(len(sext)) * 1e-3
    where = list(range(len(flat)))
    op_flags = ufo.LINEAR | ufo.ACHROMATIC | ufo.DOUBLE_PRECISION
    op = ufo.Optics(ring, where=list(where), flags=op_flags)
    co_ms, op_ms = [], []
    for obj, times in ((co, co_ms), (op, op_ms)):
        obj.run()
        for _ in range(reps):
            t = time.perf_counter()
            obj.run()
            times.append(1e3 * (time.perf_counter() - t))
    out = {"workload": f"ALBA ring ({len(flat)} elements): {len(sext)} closed orbits (one sextupole displaced "
                       f"by 1 mm each, FIVED FP64); periodic Optics at all {len(flat)} elements (FP64)",
           "closed_orbit_ms": round(float(np.median(co_ms)), 3),
           "newton_steps_max": int(np.max(co.steps)),
           "optics_ms": round(float(np.median(op_ms)), 3),
           "qx": float(op.qx), "qy": float(op.qy)}
    if cpu:
        from oracle import runner
        pool = np.array(co.parameters, dtype=np.float64)
        runner.closed_orbit(ring, pool[:1], flags=co_flags, parameters=prm)    # builds the oracle
        t = time.perf_counter()
        orbits = runner.closed_orbit(ring, pool, flags=co_flags, parameters=prm)
        out["cpu_closed_orbit_ms"] = round(1e3 * (time.perf_counter() - t), 1)
        out["closed_orbit_max_diff_vs_oracle"] = float(np.max(np.abs(orbits - co.orbits)))
        ref = runner.Optics(ring, where=list(where), flags=op_flags)
        t = time.perf_counter()
        ref.run()
        out["cpu_optics_ms"] = round(1e3 * (time.perf_counter() - t), 1)
        out["optics_max_rel_diff_beta_vs_oracle"] = float(np.max(np.abs(np.asarray(op.bx) / np.asarray(ref.bx) - 1.0)))
        out["cpu_threads"] = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        out["cpu_kind"] = "port (oracle, whole workload, build excluded)"
    return out


def batched_timing(seeds=256, turns=2048, reps=2, cpu_seconds=0.0):
    """Error-seed scan in one launch sequence: the 4-period DBA ring with
    256 seeds of misalignments (50 um) and b3-b6/a3-a6 errors (1e-4 at 10 mm),
    particle counts per seed uniform in [4096, 32768] on the DA grid law,
    2048 turns in hipGraph-captured blocks (block length and occupancy from
    the autotuner).  Timed from the graph launch to its completion.  With
    ``cpu_seconds``: the oracle on a sample of seeds 0-1, every seed's errors
    bound per particle as the reference tracks lattice variants, and the
    whole workload's host time extrapolated from its rate."""
    import numpy as np
    import torch
    import ufo_amd as ufo
    from ufo_amd import lattices
    from ufo_amd.batch import BatchedAperture
    ring, prm, vals, counts = lattices.dba_error_model(periods=4, seeds=seeds)
    ba = BatchedAperture(ring, ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED, turns, prm, vals, counts)
    ba.parameters[:] = np.concatenate([lattices.seed_grid(int(c)) for c in counts])
    ba.run()
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(ba._stream)
        ba.launch()
        e1.record(ba._stream)
        ba._stream.synchronize()
        times.append(e0.elapsed_time(e1) / 1e3)
    lost = ba._buffers()["lost"][:ba.particles].cpu().numpy()
    pt = int(np.minimum(lost.astype(np.int64) + 1, turns).sum())
    dt = float(np.mean(times))
    out = {"workload": f"{seeds} error seeds of the 4-period DBA ring ({ba.kernel.elements} elements), "
                       f"{ba.particles} particles (per-seed counts in [4096, 32768]), {turns} turns",
           "seconds": dt, "pet_per_s": pt * ba.kernel.elements / dt,
           "tflops": pt * ba.kernel.flops_per_turn / dt / 1e12,
           "flop_per_particle_turn": ba.kernel.flops_per_turn, "executed_particle_turns": pt,
           "survivors": int((lost == turns).sum()), "block_turns": ba.block_turns,
           "waves_per_simd": ba.waves_per_simd, "particles_per_lane": ba.particles_per_lane,
           "workgroup": ba.workgroup}
    if cpu_seconds:
        # the reference's own way of tracking lattice variants: one kernel with
        # every seed's errors bound per particle (HOWTO.md:20-22), seeds 0-1
        from oracle import runner
        first = int(counts[0] + counts[1])
        pool = np.concatenate([lattices.seed_grid(int(c)) for c in counts[:2]])
        pool = np.concatenate([pool, np.repeat(vals[:2], counts[:2], axis=0)], axis=1)
        flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED

        def run_rows(idx):
            return _executed(runner.stable_aperture(ring, pool[idx], flags=flags, turns=turns,
                                                    parameters=["x", "y"] + prm, mode="bench_fast"), turns)

        rate, m, secs = _cpu_sample_rate(run_rows, first, ba.kernel.elements, cpu_seconds)
        out["cpu_baseline"] = {"pet_per_s": rate, "kind": "port", "build": "bench_fast",
                               "sample": f"{m} of the {first} particles of seeds 0-1, {len(prm)} per-particle "
                                         f"parameter columns, {turns} turns, {secs:.1f} s",
                               "seconds_extrapolated": pt * ba.kernel.elements / rate,
                               "gpu_over_cpu": pt * ba.kernel.elements / dt / rate}
    return out


def seed_adjoint_timing(seeds=64, per_seed=8192, turns=1024, every=32, cpu_seconds=0.0):
    """Batched sextupole-optimisation step: the 4-period error model, 64
    seeds x 8192 particles (Gaussian sigma_x = 4 mm, sigma_y = 2 mm,
    sigma_px,py = 0.1 mrad, sigma_delta = 2e-3) launched around each seed's
    closed orbit (Newton on the tangent kernel, tolerance 1e-13 m),
    1024 turns, per-seed gradient of the per-seed amplitude/survival loss with
    respect to the 8 sextupole family strengths.  With ``cpu_seconds``: the
    oracle's rate on a sample of seed 0 (its errors bound per particle), and
    the host time of a forward pass over every seed and of the
    central-difference gradients (two passes per family) extrapolated from
    it."""
    import numpy as np
    import torch
    import ufo_amd as ufo
    from ufo_amd import lattices
    from ufo_amd.batch_adjoint import BatchedAdjoint, SeedClosedOrbit
    flags = ufo.KICK | ufo.DOUBLE_PRECISION
    ring, prm, vals, _, fam = lattices.dba_family_model(seeds=seeds)
    counts = np.full(seeds, per_seed)
    t = time.perf_counter()
    co = SeedClosedOrbit(ring, flags, prm, vals, tol=1e-13)
    co.module()
    ba = BatchedAdjoint(ring, flags, turns, ["x", "px", "y", "py", "dp"], prm, vals, counts, fam, every=every)
    ba._buffers()
    compile_s = time.perf_counter() - t
    torch.cuda.synchronize()
    t = time.perf_counter()
    orbits = co.run()                       # copies the orbits to the host: synchronous
    co_ms = 1e3 * (time.perf_counter() - t)
    rng = np.random.default_rng(4)
    n = seeds * per_seed
    x0 = np.stack([rng.normal(0, 4e-3, n), rng.normal(0, 1e-4, n), rng.normal(0, 2e-3, n),
                   rng.normal(0, 1e-4, n), rng.normal(0, 2e-3, n)], 1)
    seed_of = np.repeat(np.arange(seeds), per_seed)
    x0[:, :4] += orbits[seed_of]
    x0 = torch.tensor(x0, device=ba.device)
    seg = torch.tensor(seed_of, device=ba.device)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    fw, bw = [], []
    for rnd in range(3):                    # one warm-up round, median of two
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record(stream)
        xf, lost = ba.forward(x0)
        ev[1].record(stream)
        xf_t = xf.clone().requires_grad_(True)
        alive = torch.isfinite(xf_t[:, 0])
        r2 = torch.where(alive, xf_t[:, 0] ** 2 + xf_t[:, 2] ** 2, torch.zeros_like(xf_t[:, 0]))
        pen = torch.where(alive, torch.nn.functional.softplus(20.0 * (r2 / 0.03 ** 2 - 0.25)) / 20.0,
                          torch.ones_like(r2))
        per = torch.zeros(seeds, dtype=torch.float64, device=ba.device).index_add_(0, seg, r2 + 1e-5 * pen)
        (per / per_seed).sum().backward()
        ev[2].record(stream)
        _, gseed = ba.backward(xf_t.grad)
        ev[3].record(stream)
        torch.cuda.synchronize()
        if rnd:
            fw.append(ev[0].elapsed_time(ev[1]))
            bw.append(ev[2].elapsed_time(ev[3]))
    f_ms, b_ms = float(np.median(fw)), float(np.median(bw))
    check = seed_gradient_check(ba, ring, flags, turns, prm, vals, fam, x0, seg, seeds, per_seed)
    out = {"workload": f"{seeds} seeds x {per_seed} particles, 4-period DBA with errors, 6D, {turns} turns, "
                       f"gradient w.r.t. {len(fam)} sextupole families per seed",
           "closed_orbit_ms": round(co_ms, 2), "newton_steps_max": int(co.steps.max()),
           "seeds_converged": int(co.converged.sum()),
           "forward_ms": round(f_ms, 2), "reverse_ms": round(b_ms, 2),
           "gradient_over_forward": round((f_ms + b_ms) / f_ms, 3),
           "survivors": int((lost == turns).sum().item()), "particles": n,
           "hiprtc_compile_s": round(compile_s, 1), "gradient_finite": bool(torch.isfinite(gseed).all().item()),
           **check}
    if cpu_seconds:
        # seed 0 through the oracle with its errors bound per particle; the
        # host's route to the per-seed gradients is central differences, two
        # forward passes of every seed per family
        from oracle import runner
        E = ba.kernel.elements
        pool0 = np.concatenate([x0[:per_seed].cpu().numpy(), np.repeat(vals[:1], per_seed, axis=0)], axis=1)

        def run_rows(idx):
            return _executed(runner.stable_aperture(ring, pool0[idx], flags=flags, turns=turns,
                                                    parameters=["x", "px", "y", "py", "dp"] + prm,
                                                    mode="bench_fast"), turns)

        rate, m, secs = _cpu_sample_rate(run_rows, per_seed, E, cpu_seconds)
        pt = int(torch.clamp(lost.to(torch.int64) + 1, max=turns).sum())
        fwd_s = pt * E / rate
        out["cpu_baseline"] = {"pet_per_s": rate, "kind": "port", "build": "bench_fast",
                               "sample": f"{m} of seed 0's {per_seed} particles, {len(prm)} per-particle parameter "
                                         f"columns, {turns} turns, {secs:.1f} s",
                               "forward_seconds_extrapolated": fwd_s,
                               "central_difference_gradient_seconds_extrapolated": 2 * len(fam) * fwd_s,
                               "gpu_gradient_speedup": 2 * len(fam) * fwd_s / ((f_ms + b_ms) * 1e-3)}
    return out


def seed_gradient_check(ba, ring, flags, turns, prm, vals, fam, x0, seg, seeds, per_seed, amp_max=30.0, h=1e-3):
    """Per-seed gradients of every family against 4th-order central
    differences (steps +-h, +-2h on the family strength of every seed at
    once: seeds are independent).  The objective is the per-seed amplitude
    loss over the bunch's regular core, the survivors whose final state
    amplifies a change of the initial one by at most ``amp_max`` (from the
    adjoint's own dL/dx0): a chaotic particle's final state is no smooth
    function of the strengths at any usable step.  Reported: the largest
    |adjoint - fd| / max(|fd|, 1e-3 max_family |fd|) over seeds and
    families."""
    import numpy as np
    import torch
    from ufo_amd import batch
    from ufo_amd.batch_adjoint import BatchedAdjoint
    dev = ba.device
    xf0, _ = ba.forward(x0)
    alive = torch.isfinite(xf0[:, 0])
    lam = torch.zeros_like(xf0)
    lam[:, 0] = 2 * torch.nan_to_num(xf0[:, 0])
    lam[:, 2] = 2 * torch.nan_to_num(xf0[:, 2])
    gx0, _ = ba.backward(lam)
    mask = alive & (gx0.norm(dim=1) / lam.norm(dim=1).clamp_min(1e-300) <= amp_max)

    def per_seed_loss(xf):
        r2 = torch.where(mask, xf[:, 0] ** 2 + xf[:, 2] ** 2, torch.zeros_like(xf[:, 0]))
        soft = torch.nn.functional.softplus(20.0 * (r2 / 0.03 ** 2 - 0.25)) / 20.0
        pen = torch.where(mask, soft, torch.zeros_like(r2))
        return torch.zeros(seeds, dtype=torch.float64, device=dev).index_add_(0, seg, r2 + 1e-5 * pen) / per_seed

    xf_t = xf0.clone().requires_grad_(True)
    per_seed_loss(xf_t).sum().backward()
    g = ba.backward(torch.nan_to_num(xf_t.grad, nan=0.0))[1].cpu().numpy()
    probe = BatchedAdjoint(ring, flags, turns, ["x", "px", "y", "py", "dp"], prm, vals, np.full(seeds, per_seed),
                           fam)
    probe._buffers()

    def loss_at(v):
        probe._dev["table"] = torch.tensor(batch.build_table(probe.layout, v), dtype=torch.float64, device=dev)
        return per_seed_loss(probe.forward(x0)[0]).cpu().numpy()

    fd = np.zeros_like(g)
    for j, f in enumerate(fam):
        c = prm.index(f)
        for k, w in ((1, 8.0), (2, -1.0)):
            vp, vm = vals.copy(), vals.copy()
            vp[:, c] += k * h
            vm[:, c] -= k * h
            fd[:, j] += w * (loss_at(vp) - loss_at(vm))
        fd[:, j] /= 12.0 * h
    crit = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-3 * np.max(np.abs(fd), axis=0))
    return {"gradient_fd_max_rel": float(crit.max()),
            "gradient_fd_check": f"{seeds} seeds x {len(fam)} families vs 4th-order central differences "
                                 f"(h = {h:g}) of the loss over the regular core ({int(mask.sum())} of "
                                 f"{int(alive.sum())} survivors, amplification <= {amp_max:g})"}


def adjoint_timing(cells=48, particles=65536, turns=512, every=32, warmup=2, reps=5, cpu_seconds=0.0):
    """Gradient of the amplitude loss with respect to all 96 quadrupole k1 and
    96 sextupole k2 of a 48-cell FODO ring (65 536 particles, sigma_x =
    sigma_y = 3 mm, 512 turns, checkpoints every 32 turns): time of the
    checkpointing forward pass, of the reverse pass and of one plain forward
    pass (the StableAperture kernel of the same ring and bunch, strengths
    compiled in), HIP events on the launch stream.  Each round times the
    three back to back; ``warmup`` rounds let the clock settle, the medians
    of ``reps`` rounds are reported.  With ``cpu_seconds``: the oracle's
    forward rate on a sample of the bunch, and the host time of a forward
    pass and of the central-difference gradient (two passes per strength)
    extrapolated from it."""
    import numpy as np
    import torch
    import ufo_amd as ufo
    from ufo_amd import lattices
    ring = lattices.fodo_ring(cells, distinct=True)
    shared = ([(f"QF{i}", "k1") for i in range(cells)] + [(f"QD{i}", "k1") for i in range(cells)]
              + [(f"SF{i}", "k2") for i in range(cells)] + [(f"SD{i}", "k2") for i in range(cells)])
    theta = torch.tensor([1.2] * cells + [-1.2] * cells + [0.8] * cells + [-1.4] * cells, dtype=torch.float64)
    bunch = lattices.gaussian_bunch(particles, seed=0, sx=3e-3, sy=3e-3)
    tr = ufo.AdjointTracker(ring, ufo.KICK | ufo.DOUBLE_PRECISION, turns, ["x", "px", "y", "py", "dp"], shared,
                            every=every)
    th = theta.to(tr.device).requires_grad_(True)
    x0 = torch.tensor(bunch, device=tr.device)
    sa = ufo.StableAperture(ring, flags=ufo.KICK | ufo.DOUBLE_PRECISION, turns=turns, particles=particles,
                            parameters=["x", "px", "y", "py", "dp"])
    sa.parameters[:] = bunch
    sa.module()
    sa.pool_device.copy_(sa._pool_host)
    stream = torch.cuda.current_stream()
    fw, bw, pf = [], [], []
    for k in range(warmup + reps):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        e[0].record(stream)
        xf, lost = tr.forward(x0, th.detach())
        e[1].record(stream)
        lam = torch.zeros_like(xf)
        alive = torch.isfinite(xf[:, 0])
        lam[alive] = 2.0 * xf[alive] * torch.tensor([1.0, 0.0, 1.0, 0.0], device=tr.device) / particles
        e[2].record(stream)
        tr.backward(lam)
        e[3].record(stream)
        sa.launch()
        e[4].record(stream)
        torch.cuda.synchronize()
        if k >= warmup:
            fw.append(e[0].elapsed_time(e[1]))
            bw.append(e[2].elapsed_time(e[3]))
            pf.append(e[3].elapsed_time(e[4]))
    f_ms, b_ms, p_ms = float(np.median(fw)), float(np.median(bw)), float(np.median(pf))
    pt = int(torch.clamp(lost.to(torch.int64) + 1, max=turns).sum())
    flop = pt * tr.kfwd.flops_per_turn
    out = {"workload": f"{cells}-cell FODO, {particles} particles, {turns} turns, checkpoints every {every}, "
                       f"gradient w.r.t. {len(shared)} strengths",
           "forward_ms": f_ms, "reverse_ms": b_ms, "gradient_over_forward": (f_ms + b_ms) / f_ms,
           "plain_forward_ms": p_ms, "gradient_over_plain_forward": (f_ms + b_ms) / p_ms,
           "forward_tflops": flop / (f_ms * 1e-3) / 1e12, "survivors": int((lost == turns).sum()),
           "rounds": reps}
    if cpu_seconds:
        # the host has no adjoint: its route to the gradient is central
        # differences, two forward passes per strength
        from oracle import runner
        E = sa.kernel.elements

        def run_rows(idx):
            return _executed(runner.stable_aperture(ring, bunch[idx], flags=ufo.KICK | ufo.DOUBLE_PRECISION,
                                                    turns=turns, parameters=["x", "px", "y", "py", "dp"],
                                                    mode="bench_fast"), turns)

        rate, m, secs = _cpu_sample_rate(run_rows, particles, E, cpu_seconds)
        fwd_s = pt * E / rate
        out["cpu_baseline"] = {"pet_per_s": rate, "kind": "port", "build": "bench_fast",
                               "sample": f"{m} of the {particles} particles, {turns} turns, {secs:.1f} s",
                               "forward_seconds_extrapolated": fwd_s,
                               "central_difference_gradient_seconds_extrapolated": 2 * len(shared) * fwd_s,
                               "gpu_gradient_speedup": 2 * len(shared) * fwd_s / ((f_ms + b_ms) * 1e-3)}
    return out


def measured_traffic(args):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (copy-calibrated), when the run is that
    workload; None otherwise."""
    if (args.nx, args.ny, args.turns, args.sext_scale) != (1024, 1024, 1024, 1.0):
        return None
    path = os.path.join(ROOT, "profiles", "ufo_sa_dba_da_scan_1M_1024turns.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return int(line.split("=")[1])
    except OSError:
        return None
    return None


def _host_record():
    """The host the CPU baseline ran on: lscpu's model name (else
    /proc/cpuinfo's), logical CPUs (nproc), CPUs in this process's affinity
    mask, and physical cores when lscpu reports them."""
    import subprocess
    rec = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "model": None,
           "physical_cores": None}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        fields = dict(line.split(":", 1) for line in text.splitlines() if ":" in line)
        fields = {k.strip(): v.strip() for k, v in fields.items()}
        rec["model"] = fields.get("Model name")
        cores, sockets = fields.get("Core(s) per socket"), fields.get("Socket(s)")
        if cores and sockets and cores.isdigit() and sockets.isdigit():
            rec["physical_cores"] = int(cores) * int(sockets)
    except (OSError, subprocess.SubprocessError, ValueError):
        pass
    if not rec["model"]:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        rec["model"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return rec


def _cpu_sample_rate(run_rows, n, E, budget_s):
    """Host rate of the oracle over an evenly strided sample of rows
    0..n-1 sized to ~budget_s seconds: (PET/s, sample size, seconds).
    ``run_rows(indices)`` runs the oracle (gcc fast-math build, OpenMP over
    the host threads) on those rows and returns their executed
    particle-turns."""
    import numpy as np
    rows = np.arange(n)
    probe = rows[:: max(1, n // 128)][:128]
    run_rows(probe)                              # builds (cached) and warms the library
    t = time.perf_counter()
    run_rows(probe)
    per_row = (time.perf_counter() - t) / len(probe)
    m = int(min(n, max(128, budget_s / max(per_row, 1e-9))))
    sample = rows[:: max(1, n // m)][:m]
    t = time.perf_counter()
    pt = run_rows(sample)
    dt = time.perf_counter() - t
    return pt * E / dt, len(sample), dt


def _executed(lost, turns):
    import numpy as np
    return int(np.minimum(np.asarray(lost).astype(np.int64) + 1, turns).sum())


def cpu_baseline(ring, flags, pool, turns, E, budget_s, gpu_lost=None):
    """The oracle (restated reference kernels, OpenMP over the host threads)
    on an evenly strided sample of the grid sized to ~budget_s seconds, in
    two builds: gcc -O3 -march=native fast-math (the analogue of the
    reference's default -cl-fast-relaxed-math -cl-mad-enable, NaN tests
    kept; ``value``) and strict IEEE (-ffp-contract=off, on a sample sized
    to half the budget; ``strict``).  With ``gpu_lost`` (the GPU's loss
    turns over the whole grid) it also reports on how many sample points
    the fast build's loss turns are identical to the GPU's."""
    import numpy as np
    from oracle import runner
    host = _host_record()
    threads = int(os.environ.get("OMP_NUM_THREADS") or host["affinity_cpus"] or 1)
    res = {}
    for mode, budget in (("bench_fast", budget_s), ("bench_strict", budget_s / 2)):
        probe = pool[:: max(1, len(pool) // 256)][:256]
        runner.stable_aperture(ring, probe, flags=flags, turns=turns, parameters=["x", "y"], mode=mode)
        t = time.perf_counter()
        runner.stable_aperture(ring, probe, flags=flags, turns=turns, parameters=["x", "y"], mode=mode)
        per_particle = (time.perf_counter() - t) / len(probe)
        m = int(min(len(pool), max(256, budget / max(per_particle, 1e-9))))
        stride = max(1, len(pool) // m)
        sample = pool[::stride][:m]
        t = time.perf_counter()
        lost = runner.stable_aperture(ring, sample, flags=flags, turns=turns, parameters=["x", "y"], mode=mode)
        dt = time.perf_counter() - t
        pt = int(np.minimum(lost.astype(np.int64) + 1, turns).sum())
        res[mode] = (pt * E / dt, stride, sample, lost, dt)
    value, stride, sample, lost, dt = res["bench_fast"]
    s_value, s_stride, s_sample, _, s_dt = res["bench_strict"]
    out = {"value": value, "unit": "particle-element-turns/s", "cores": threads, "kind": "port",
           "sample": f"{len(sample)} grid points (one in {stride} of the {len(pool)}-point grid), "
                     f"{turns} turns, {dt:.1f} s",
           "build": "gcc " + " ".join(runner.FLAGS["bench_fast"]) + " -fopenmp",
           "strict": {"value": s_value, "build": "gcc " + " ".join(runner.FLAGS["bench_strict"]) + " -fopenmp",
                      "sample": f"{len(s_sample)} grid points (one in {s_stride}), {turns} turns, {s_dt:.1f} s"},
           "host": host}
    if gpu_lost is not None:
        same = np.asarray(gpu_lost)[::stride][:len(sample)] == lost
        out["lost_identical"] = f"{int(same.sum())}/{len(sample)}"
        out["same_gpu_reference_opencl"] = reference_opencl_scan(ring, flags, pool, turns, E, gpu_lost)
    return out


def reference_opencl_scan(ring, flags, pool, turns, E, gpu_lost, threads=524288):
    """The reference's own StableAperture kernel (its templates in the
    kernel body of ufo/sa.py, built by the device's OpenCL compiler with its
    default options -cl-fast-relaxed-math -cl-mad-enable, ``threads``
    work-items claiming particles with atomic_inc) on the same GPU over the
    whole workload: kernel time (OpenCL profiling events, best of 2) and
    whether its loss turns equal the HIP kernel's.  None without an OpenCL
    GPU device."""
    import numpy as np
    from oracle import opencl
    if not opencl.available():
        return None
    threads = min(threads, len(pool))
    lost, ms = opencl.stable_aperture(ring, pool, flags, turns, ["x", "y"], threads=threads, reps=2)
    pt = int(np.minimum(lost.astype(np.int64) + 1, turns).sum())
    return {"kernel_ms": round(min(ms), 2), "pet_per_s": pt * E / (min(ms) * 1e-3), "threads": threads,
            "device": opencl.device().name, "options": opencl.DEFAULT_CL_OPTIONS,
            "lost_identical_to_hip": f"{int(np.sum(lost == np.asarray(gpu_lost)))}/{len(pool)}"}


if __name__ == "__main__":
    main()
