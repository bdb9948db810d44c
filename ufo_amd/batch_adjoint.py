"""Per-seed closed orbits and per-seed gradients for many lattice variants.

``SeedClosedOrbit``: one closed orbit per error seed by Newton iteration on
the one-turn map (forward-tangent Jacobian), each lane one seed reading its
own coefficient table; dp = 0 as the reference's ClosedOrbit objective.

``BatchedAdjoint``: forward tracking with checkpoints and the reverse pass
for many seeds in one launch each.  Particles are grouped by seed in tiles of
64 (one wave, one seed); the per-seed lattice parameters come from the
per-seed coefficient table; the gradient of a final-state loss with respect
to a chosen subset of the per-seed parameter columns (e.g. sextupole family
strengths) is accumulated per wave in LDS and added to that seed's row of
the output.  Kicks whose strength carries 1/(1+dp) with a per-particle dp
are computed without the factor and scaled as a whole.
"""
from __future__ import annotations

from typing import List

import numpy as np
import torch

from . import adjoint, batch, codegen, ir, tangent
from .lattice import lower
from .adjoint import ACCW
from .flags import DOUBLE_PRECISION
from .sim import Simulation

TILE = 64


def _ops(line, flags, parameters):
    return lower(line, flags, parameters)


def _table_env(flags, src, ncoord, diff_cols=None):
    env = codegen._env_for(flags, src, None)
    env.shared_offset = ncoord
    env.table = []
    env.period_slots = {}
    env.no_loop = True
    env.diff_cols = diff_cols
    return env


def _layout_of(env, ncoord):
    """Table layout: flat slots, then the period slots of every repetition of
    the period loop (if the turn has one), with the parameter stride or the
    per-repetition column maps that turn repetition 0's slot expressions
    into repetition r's."""
    head, p, reps, stride = env.last_period
    F, S = len(env.table), len(env.period_slots)
    reps = int(reps) if reps >= 2 and S else 0
    return {"flat": [e for e, _ in env.table], "period": sorted(env.period_slots, key=env.period_slots.get),
            "reps": reps, "stride": stride if isinstance(stride, list) else int(stride), "F": F, "S": S,
            "nslots": max(1, F + reps * S), "offset": ncoord}


def _fill(text, layout):
    return text.replace("@PB@", str(layout["F"])).replace("@PS@", str(max(layout["S"], 1)))


def co_seed_kernel(line, flags, lattice_parameters, iterations=32, tol=1e-13, name="ufo_co_seed"):
    """Newton closed orbit per seed (lane = seed), dp = 0 (ufo/co.py:175-193).
    ``conv[sd]`` is 1 only when the residual reached ``tol``; an orbit that
    is lost or does not reach it within ``iterations`` steps is returned as
    the evaluated iterate with the smallest residual."""
    coords: List[str] = []
    ops = _ops(line, flags, coords + list(lattice_parameters))
    src = codegen._coord_sources(coords, 0, 0.0)
    env = _table_env(flags, src, 0)
    env.no_loop = False
    env.remap_loops = True
    body, _ = codegen._layout(ops, env, emit=tangent.emit_with_tangent)
    layout = _layout_of(env, 0)
    body = [_fill(b, layout) for b in body]
    one, zero = env.lit(1.0), env.lit(0.0)
    code = codegen._PRELUDE.format(ufloat="double", nan='__builtin_nan("")') + "#define NT 4\n" + tangent._SOLVE4
    code += f"""
extern "C" __global__ void __launch_bounds__(64)
{name}(const ufloat *__restrict__ table, const ufloat *__restrict__ guess, ufloat *__restrict__ orbits,
       int *__restrict__ steps, int *__restrict__ conv, int seeds)
{{
    {codegen._decls(env)}
    ufloat tx[NT], tpx[NT], ty[NT], tpy[NT];
    const int sd = blockIdx.x * 64 + threadIdx.x;
    if (sd >= seeds) return;
    const ufloat *__restrict__ ct = table + (size_t)sd * {layout['nslots']};
    z = {zero}; dp = {zero}; oodppo = {one};
{chr(10).join('    ' + c for c in env.inv_code)}
    ufloat X[4] = {{guess[sd * 4], guess[sd * 4 + 1], guess[sd * 4 + 2], guess[sd * 4 + 3]}};
    ufloat best = __builtin_huge_val(), XB[4] = {{X[0], X[1], X[2], X[3]}};
    bool done = false;
    int it = 0;
    for (; it < {int(iterations)}; ++it) {{
        x = X[0]; px = X[1]; y = X[2]; py = X[3];
#pragma unroll
        for (int k = 0; k < NT; ++k) {{
            tx[k] = (k == 0) ? {one} : {zero}; tpx[k] = (k == 1) ? {one} : {zero};
            ty[k] = (k == 2) ? {one} : {zero}; tpy[k] = (k == 3) ? {one} : {zero};
        }}
        {{
            const int turn = 0;
{chr(10).join('            ' + c for c in body)}
        }}
        ufloat r[4] = {{x - X[0], px - X[1], y - X[2], py - X[3]}};
        if (isnan(x) || isnan(px) || isnan(y) || isnan(py)) break;
        ufloat rmax = fmax(fmax(fabs(r[0]), fabs(r[1])), fmax(fabs(r[2]), fabs(r[3])));
        if (rmax <= {env.lit(tol)}) {{ done = true; break; }}
        if (rmax < best) {{ best = rmax; XB[0] = X[0]; XB[1] = X[1]; XB[2] = X[2]; XB[3] = X[3]; }}
        ufloat A[4][4];
#pragma unroll
        for (int k = 0; k < NT; ++k) {{
            A[0][k] = tx[k]; A[1][k] = tpx[k]; A[2][k] = ty[k]; A[3][k] = tpy[k];
        }}
#pragma unroll
        for (int i = 0; i < 4; ++i) A[i][i] -= {one};
        ufo_solve4(A, r);
#pragma unroll
        for (int i = 0; i < 4; ++i) X[i] -= r[i];
    }}
    if (!done) {{ X[0] = XB[0]; X[1] = XB[1]; X[2] = XB[2]; X[3] = XB[3]; }}
    ufloat *o = orbits + (size_t)sd * 4;
    o[0] = X[0]; o[1] = X[1]; o[2] = X[2]; o[3] = X[3];
    steps[sd] = it;
    conv[sd] = done ? 1 : 0;
}}
"""
    return codegen.Kernel(name=name, src=code, kind="co_seed", fp32=False, flops_per_turn=codegen._count(ops, env),
                          elements=codegen.element_count(line), block=64), layout


class SeedClosedOrbit(Simulation):
    """Closed orbit of every seed (x, px, y, py at the ring start, dp = 0).
    ``seed_values`` [seeds, len(lattice_parameters)]; ``guess`` [seeds, 4]
    (default 0).  After ``run()``: ``orbits`` [seeds, 4], ``steps`` (Newton
    steps taken) and ``converged`` (the residual reached ``tol``; a seed whose
    orbit is lost or stays above ``tol`` gets the evaluated iterate with the
    smallest residual and False)."""

    def __init__(self, line, flags, lattice_parameters, seed_values, iterations=32, tol=1e-13, context=None,
                 options=None):
        self._init_common(flags | DOUBLE_PRECISION, context, options)
        self.kernel, self.layout = co_seed_kernel(line, flags | DOUBLE_PRECISION, lattice_parameters, iterations, tol)
        self.seed_values = np.asarray(seed_values, dtype=np.float64)
        self.seeds = self.seed_values.shape[0]
        self.table = batch.build_table(self.layout, self.seed_values)
        self.guess = np.zeros((self.seeds, 4))
        self.iterations = int(iterations)

    def run(self, threads=1):
        d = self.device
        tab = torch.tensor(self.table, dtype=torch.float64, device=d)
        guess = torch.tensor(self.guess, dtype=torch.float64, device=d)
        orbits = torch.empty((self.seeds, 4), dtype=torch.float64, device=d)
        steps = torch.empty((self.seeds,), dtype=torch.int32, device=d)
        conv = torch.empty((self.seeds,), dtype=torch.int32, device=d)
        if self.seeds:
            self.module().launch(self.kernel.name, (self.seeds + 63) // 64, 64,
                                 [tab, guess, orbits, steps, conv, self.seeds])
        self.orbits = orbits.cpu().numpy()
        self.steps = steps.cpu().numpy()
        self.converged = conv.cpu().numpy() != 0
        return self.orbits


def _reverse_turn(ops, env, diff_cols, period):
    """Reverse pass of one turn for the batched adjoint: the ops backwards,
    the repeated period as a descending loop (as the forward kernel's).
    Gradient sums of a segment are flushed into the per-seed LDS
    accumulators; inside the loop the accumulator of a sum depends on the
    repetition (the parameters of repetition r are repetition 0's mapped by
    the period's stride or column maps), looked up in a constant table
    ``ufo_gmap[rep][k]``.  Returns (lines, gmap rows or None)."""
    head, p, reps, stride = period
    looped = reps >= 2

    def image(i, r):
        if isinstance(stride, list):
            return stride[r].get(i, i)
        return i + r * int(stride)

    gmap_rows = [[] for _ in range(reps)] if looped else []

    def segment(seg_ops, in_loop):
        g = adjoint._Grad()
        body: List[str] = []
        for op in reversed(adjoint.merge_drifts(seg_ops, env.f)):
            adjoint.emit_reverse(op, env, body, g)
        decl = [f"ufloat {r} = 0.0;" for r in g.sums.values()]
        flush = []
        for i in g.params():
            terms = [(r, dd) for (ii, dd), r in g.sums.items() if ii == i]
            acc_ = None
            for r, dd in terms:
                dv = env.val(dd)
                acc_ = f"{r} * {dv}" if acc_ is None else f"fma({r}, {dv}, {acc_})"
            if in_loop:
                col = len(gmap_rows[0])
                for r in range(reps):
                    c = image(i, r)
                    if c not in diff_cols:
                        raise codegen.NeedsFlat("a repetition maps a gradient onto an undifferentiated column")
                    gmap_rows[r].append(diff_cols.index(c))
                idx = f"ufo_gmap[rep][{col}]"
            else:
                idx = str(diff_cols.index(i))
            flush.append(f"unsafeAtomicAdd(&acc[{idx} * {ACCW} + slot], {acc_});")
        return ["{"] + ["    " + x for x in decl + body + flush] + ["}"]

    out: List[str] = []
    if looped:
        import copy
        saved = copy.deepcopy((env.table, env.period_slots, env.invariants, env.inv_code, env.tmp))
        try:
            tail = segment(ops[head + p * reps:], False)
            env.stride = ("rep", stride)
            try:
                inner = segment(ops[head:head + p], True)
            finally:
                env.stride = None
            out += tail
            out.append("#pragma clang loop unroll(disable)")
            out.append(f"for (int rep = {reps - 1}; rep >= 0; --rep) {{")
            out += ["    " + x for x in inner]
            out.append("}")
            out += segment(ops[:head], False)
            return out, (gmap_rows if gmap_rows and gmap_rows[0] else None)
        except codegen.NeedsFlat:
            # restore the shared table and emit the turn flat
            tb, ps, inv, code, tmp = saved
            env.table[:] = tb
            env.period_slots.clear()
            env.period_slots.update(ps)
            env.invariants, env.inv_code, env.tmp = inv, code, tmp
            out = []
    return segment(ops, False), None


def batched_adjoint_kernels(line, flags, turns, coords, lattice_parameters, diff_parameters, dp=0.0, every=32):
    """Forward (checkpointing) and reverse kernels of the batched adjoint;
    ``diff_parameters`` is the subset of ``lattice_parameters`` whose
    gradients are accumulated, per seed."""
    coords = list(coords)
    lp = list(lattice_parameters)
    ops = _ops(line, flags, coords + lp)
    for op in ops:
        if isinstance(op, ir.Code) or any(e.has_raw() or e.uses_var("turn") for e in op.exprs()):
            raise NotImplementedError("raw user code has no adjoint")
    off = len(coords)
    diff_cols = [off + lp.index(p) for p in diff_parameters]
    nck = (turns + every - 1) // every + 1
    src = codegen._coord_sources(coords, flags, dp)
    P = max(1, len(coords))
    nd = max(1, len(diff_cols))
    pre = codegen._PRELUDE.format(ufloat="double", nan='__builtin_nan("")')

    # one table for both kernels: the reverse env continues the forward's
    envf = _table_env(flags, src, off)
    envf.no_loop = False
    envf.remap_loops = True
    fbody, _ = codegen._layout(ops, envf)
    envb = _table_env(flags, src, off, diff_cols)
    envb.table, envb.period_slots = envf.table, envf.period_slots
    rbody, gmaps = _reverse_turn(ops, envb, diff_cols, envf.last_period)
    layout = _layout_of(envf, off)
    nslots = layout["nslots"]
    fbody = [_fill(x, layout) for x in fbody]
    rbody = [_fill(x, layout) for x in rbody]
    inv_b = [_fill(x, layout) for x in envb.inv_code]
    gtab = ""
    if gmaps:
        rows = ", ".join("{" + ", ".join(str(v) for v in row) + "}" for row in gmaps)
        gtab = f"static __constant__ int ufo_gmap[{len(gmaps)}][{len(gmaps[0])}] = {{{rows}}};\n"

    def load(env, inv, indent):
        lines = []
        for c in codegen.COORDS:
            kind, v = src[c]
            lines.append(f"{c} = row[{v}];" if kind == "p" else f"{c} = {env.lit(v)};")
        if "oodppo" in env.consts:
            lines.append(f"oodppo = {env.lit(env.consts['oodppo'])};")
        else:
            lines.append("oodppo = 1.0 / (dp + 1.0);")
        lines += inv
        return [indent + x for x in lines]

    tile_head = f"""
    const int tile = blockIdx.x;
    const int sd = tile_seed[tile];
    const int lane = threadIdx.x;
    const int idx = tile_start[tile] + lane;
    const bool active = lane < tile_count[tile];
    const size_t np = (size_t)particles;
    const ufloat *__restrict__ ct = table + (size_t)sd * {nslots};
"""
    store_ck = ("{ ufloat *c = ckpt + (size_t)ck * 4 * np + idx; c[0] = x; c[np] = px; c[2 * np] = y; "
                "c[3 * np] = py; }")
    fwd = pre + f"""
extern "C" __global__ void __launch_bounds__(64)
ufo_badj_fwd(const ufloat *__restrict__ pool, const ufloat *__restrict__ table, const int *__restrict__ tile_seed,
             const int *__restrict__ tile_start, const int *__restrict__ tile_count, ufloat *__restrict__ ckpt,
             ufloat *__restrict__ xout, int *__restrict__ lost, int particles)
{{
    {codegen._decls(envf)}
{tile_head}
    if (!active) return;
    const ufloat *__restrict__ row = pool + (size_t)idx * {P};
{chr(10).join(load(envf, envf.inv_code, '    '))}
    int ck = 0;
    {store_ck}
    int turn = 0;
    for (; turn < {turns}; ++turn) {{
{chr(10).join('        ' + s_ for s_ in fbody)}
        if (isnan(x)) break;
        if ((turn + 1) % {every} == 0 || turn + 1 == {turns}) {{ ++ck; {store_ck} }}
    }}
    ufloat *o = xout + (size_t)idx * 4;
    o[0] = x; o[1] = px; o[2] = y; o[3] = py;
    lost[idx] = turn;
}}
"""
    bwd = pre + gtab + f"""
extern "C" __global__ void __launch_bounds__(64)
ufo_badj_bwd(const ufloat *__restrict__ pool, const ufloat *__restrict__ table, const int *__restrict__ tile_seed,
             const int *__restrict__ tile_start, const int *__restrict__ tile_count, const ufloat *__restrict__ ckpt,
             const ufloat *__restrict__ lam, const int *__restrict__ lost, ufloat *__restrict__ gx0,
             ufloat *__restrict__ gseed, int particles)
{{
    __shared__ ufloat acc[{nd} * {ACCW}];
    for (int i = threadIdx.x; i < {nd} * {ACCW}; i += 64) acc[i] = 0.0;
    __syncthreads();
{tile_head}
    const int slot = lane % {ACCW};
    if (active && lost[idx] == {turns}) {{
        {codegen._decls(envb)}
        const ufloat *__restrict__ row = pool + (size_t)idx * {P};
{chr(10).join(load(envb, inv_b, '        '))}
        ufloat lx = lam[(size_t)idx * 4], lpx = lam[(size_t)idx * 4 + 1];
        ufloat ly = lam[(size_t)idx * 4 + 2], lpy = lam[(size_t)idx * 4 + 3];
        for (int ck = {nck - 1}; ck >= 1; --ck) {{
            const int t_end = (ck * {every} < {turns}) ? ck * {every} : {turns};
            const int t_begin = (ck - 1) * {every};
            const ufloat *c = ckpt + (size_t)ck * 4 * np + idx;
            x = c[0]; px = c[np]; y = c[2 * np]; py = c[3 * np];
            for (int turn = t_end - 1; turn >= t_begin; --turn) {{
{chr(10).join('                ' + s_ for s_ in rbody)}
            }}
        }}
        ufloat *o = gx0 + (size_t)idx * 4;
        o[0] = lx; o[1] = lpx; o[2] = ly; o[3] = lpy;
    }}
    __syncthreads();
    for (int j = threadIdx.x; j < {nd}; j += 64) {{
        ufloat s = 0.0;
        for (int k = 0; k < {ACCW}; ++k) s += acc[j * {ACCW} + k];
        unsafeAtomicAdd(&gseed[(size_t)sd * {nd} + j], s);
    }}
}}
"""
    fk = codegen.Kernel(name="ufo_badj_fwd", src=fwd, kind="badj_fwd", fp32=False,
                        flops_per_turn=codegen._count(ops, envf), elements=codegen.element_count(line), block=64)
    bk = codegen.Kernel(name="ufo_badj_bwd", src=bwd, kind="badj_bwd", fp32=False,
                        flops_per_turn=codegen._count(ops, envf), elements=codegen.element_count(line), block=64)
    return fk, bk, layout, nck


class BatchedAdjoint(Simulation):
    """Forward tracking and per-seed gradients for many lattice variants.

    ``lattice_parameters``: per-seed parameter columns (tuples, the
    reference's convention), ``seed_values`` [seeds, P]; ``counts``
    particles per seed (rows of x0 grouped by seed); ``diff_parameters``:
    the subset of ``lattice_parameters`` to differentiate.
    ``forward(x0)`` -> (final state [N, 4], lost turn [N]);
    ``backward(lam)`` -> (dL/dx0 [N, 4], dL/dtheta per seed [seeds, D])."""

    def __init__(self, line, flags, turns, coords, lattice_parameters, seed_values, counts, diff_parameters,
                 dp=0.0, every=32, context=None, options=None):
        self._init_common(flags | DOUBLE_PRECISION, context, options)
        self.turns = int(turns)
        self.coords = list(coords)
        self.kfwd, self.kbwd, self.layout, self.ncheck = batched_adjoint_kernels(
            line, flags | DOUBLE_PRECISION, self.turns, self.coords, lattice_parameters, diff_parameters, dp, every)
        self.kernel = self.kfwd
        self.counts = np.asarray(counts, dtype=np.int64).reshape(-1)
        if len(self.counts) == 0 or np.any(self.counts < 0):
            raise ValueError("counts: one non-negative particle count per seed, at least one seed")
        self.seed_values = np.asarray(seed_values, dtype=np.float64).reshape(len(self.counts), -1)
        if self.seed_values.shape[1] != len(lattice_parameters):
            raise ValueError(f"seed_values has {self.seed_values.shape[1]} columns for "
                             f"{len(lattice_parameters)} lattice parameters")
        self.table = batch.build_table(self.layout, self.seed_values)
        self.ndiff = max(1, len(diff_parameters))
        seeds, starts, cnts = [], [], []
        off = 0
        for s, c in enumerate(self.counts):
            for t0 in range(0, int(c), TILE):
                seeds.append(s)
                starts.append(off + t0)
                cnts.append(min(TILE, int(c) - t0))
            off += int(c)
        self.particles = off
        self._tiles = (np.array(seeds, np.int32), np.array(starts, np.int32), np.array(cnts, np.int32))
        self._dev = None
        self._saved = None

    def _buffers(self):
        if self._dev is None:
            d = self.device
            self._dev = {"table": torch.tensor(self.table, dtype=torch.float64, device=d),
                         "ts": torch.tensor(self._tiles[0], device=d), "t0": torch.tensor(self._tiles[1], device=d),
                         "tc": torch.tensor(self._tiles[2], device=d)}
            self._mods = (self.ctx.module(self.kfwd.src, self.options), self.ctx.module(self.kbwd.src, self.options))
        return self._dev

    def forward(self, x0):
        """x0 [particles, len(coords)], rows grouped by seed.  The buffers
        the reverse pass needs are kept for the next ``backward()`` call
        without ``saved``."""
        xf, lost, self._saved = self.forward_saved(x0)
        return xf, lost

    def forward_saved(self, x0):
        """As ``forward``, returning (final state, lost, saved) where
        ``saved`` (x0, checkpoints, lost) belongs to this pass only: give it
        to ``backward(lam, saved)`` to differentiate this pass even after
        later forwards."""
        b = self._buffers()
        x0 = torch.as_tensor(x0, dtype=torch.float64, device=self.device).contiguous()
        n = self.particles
        if tuple(x0.shape) != (n, len(self.coords)):
            raise ValueError(f"x0 has shape {tuple(x0.shape)}; expected ({n}, {len(self.coords)})")
        ckpt = torch.empty((self.ncheck, 4, max(1, n)), dtype=torch.float64, device=self.device)
        xf = torch.empty((n, 4), dtype=torch.float64, device=self.device)
        lost = torch.empty((n,), dtype=torch.int32, device=self.device)
        nt = len(self._tiles[0])
        if nt:
            self._mods[0].launch(self.kfwd.name, nt, 64,
                                 [x0, b["table"], b["ts"], b["t0"], b["tc"], ckpt, xf, lost, n])
        return xf, lost, (x0, ckpt, lost)

    def backward(self, lam, saved=None):
        """lam = dL/d(final x, px, y, py) [particles, 4] (NaN rows of lost
        particles count as 0) for the pass whose ``saved`` buffers are given
        (default: the last ``forward``)."""
        b = self._buffers()
        if saved is None:
            saved = self._saved
        if saved is None:
            raise RuntimeError("backward() needs a forward() first")
        x0, ckpt, lost = saved
        n = self.particles
        lam = torch.nan_to_num(torch.as_tensor(lam, dtype=torch.float64, device=self.device), nan=0.0).contiguous()
        if tuple(lam.shape) != (n, 4):
            raise ValueError(f"lam has shape {tuple(lam.shape)}; expected ({n}, 4)")
        gx0 = torch.zeros((n, 4), dtype=torch.float64, device=self.device)
        gseed = torch.zeros((len(self.counts), self.ndiff), dtype=torch.float64, device=self.device)
        nt = len(self._tiles[0])
        if nt:
            self._mods[1].launch(self.kbwd.name, nt, 64,
                                 [x0, b["table"], b["ts"], b["t0"], b["tc"], ckpt, lam, lost, gx0, gseed, n])
        return gx0, gseed
