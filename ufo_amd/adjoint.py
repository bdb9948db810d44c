"""Reverse-mode adjoint of the tracker: gradients of a final-state loss with
respect to shared lattice parameters (quadrupole k1, sextupole and multipole
strengths, misalignments, lengths) and to the initial coordinates.

Forward kernel ``ufo_adj_fwd``: tracks every particle ``turns`` turns with
the lattice parameters read from a shared vector theta (wave-uniform scalar
loads), stores the state every ``every`` turns (checkpoints, structure of
arrays [checkpoint][coordinate][particle]), the final state and the loss
turn.

Reverse kernel ``ufo_adj_bwd``: per surviving particle, from lambda =
dL/d(final state), walks the checkpoint blocks backwards.  Inside a block the
intermediate states are recomputed by the exact inverse of every op (drift:
x -= px L; kick: px += Re w, py -= Im w; thick: the inverse matrix), starting
from the block's end checkpoint; each block starts again from its stored
checkpoint, so reconstruction errors never build up over more than
``every`` turns.  Per op, in reverse order: restore the pre-op state,
accumulate the parameter gradient (the op's coefficient adjoint times the
symbolic derivative of the coefficient with respect to the parameter), then
lambda <- J^T lambda.

Gradient accumulation: contributions of one parameter within one period are
summed in registers and flushed once per period iteration into per-wave LDS
accumulators (16 slots per parameter, lane % 16, hardware FP64 LDS atomics);
at the end each block writes its per-parameter partial sums.  Lost particles
(the loss ignores them) do not enter the reverse pass.
"""
from __future__ import annotations

import math
from typing import Dict, List

import torch

from . import codegen, ir
from .lattice import lower
from .codegen import Env, _fma, _fma_neg, is_zero, kick_coeffs
from .flags import DOUBLE_PRECISION
from .sim import Simulation

ACCW = 16   # LDS accumulator slots per parameter (lane % ACCW)


class _Grad:
    """Gradient registers of one emitted segment.  Coefficient adjoints are
    summed per (parameter, derivative) pair — the slices of one magnet share
    both — and scaled by the derivative once, at the flush."""

    def __init__(self):
        self.sums: Dict[tuple, str] = {}

    def reg(self, i: int, d) -> str:
        key = (i, d)
        if key not in self.sums:
            self.sums[key] = f"S{len(self.sums)}"
        return self.sums[key]

    def params(self):
        return sorted({i for i, _ in self.sums})


def _add_grad(env: Env, g: _Grad, coef: ir.Expr, adj, out: List[str]):
    """Accumulate adj * d coef / d theta_i for every shared parameter i in
    coef.  ``adj`` is a C expression, or a list of (factor, variable) pairs
    summed with fused multiply-adds."""
    for i in sorted(ir.Expr.of(coef).params()):
        if i < env.shared_offset or (env.diff_cols is not None and i not in env.diff_cols):
            continue
        d = env.f(ir.diff(coef, i))
        if is_zero(d):
            continue
        r = g.reg(i, d)
        if isinstance(adj, list):
            for a, b in adj:
                out.append(f"{r} = fma({a}, {b}, {r});")
        else:
            out.append(f"{r} = {r} + {adj};")


def emit_reverse(op, env: Env, out: List[str], g: _Grad):
    f = env.f
    if isinstance(op, ir.Drift):
        L = f(op.L)
        if is_zero(L):
            return
        # with scaled momenta p' = p (1 + dp) the drift is x += p' L/(1+dp)
        coef = op.L * ir.OODPPO if env.scaled_momenta else op.L
        v = env.val(coef)
        out.append(f"x = fma(-px, {v}, x); y = fma(-py, {v}, y);")
        _add_grad(env, g, coef, [("lx", "px"), ("ly", "py")], out)
        out.append(f"lpx = fma(lx, {v}, lpx); lpy = fma(ly, {v}, lpy);")
    elif isinstance(op, ir.ExactDrift):
        L = env.val(op.L)
        opdp2 = env.val(ir.E(1.0 + ir.DP) * (1.0 + ir.DP))
        one = env.lit(1.0)
        out.append(f"{{ ufloat ia = {one} / sqrt({_fma('-py', 'py', _fma('-px', 'px', opdp2))}); "
                   f"ufloat l = {L} * ia; ufloat c = l * ia * ia; "
                   f"x = fma(-px, l, x); y = fma(-py, l, y); "
                   f"ufloat s = c * fma(lx, px, ly * py); "
                   f"lpx = fma(lx, l, fma(px, s, lpx)); lpy = fma(ly, l, fma(py, s, lpy)); }}")
        if ir.Expr.of(op.L).params():
            raise NotImplementedError("gradient with respect to an exact-drift length")
    elif isinstance(op, ir.Kick):
        if env.scaled_momenta:
            # p' = p (1 + dp) takes the kick's 1/(1+dp) factor
            if not op.chroma:
                raise ValueError("scaled momenta need every kick scaled by 1/(1+dp)")
            op = ir.Kick(op.kn, op.ks, False)
        emit_kick_reverse(op, env, out, g)
    elif isinstance(op, ir.Weak):
        w = f(op.w)
        if not is_zero(w):
            v = env.val(w)
            out.append(f"px = fma({v}, x, px);")
            _add_grad(env, g, op.w, [("-lpx", "x")], out)
            out.append(f"lx = fma({env.neg(w)}, lpx, lx);")
    elif isinstance(op, ir.Shift):
        dx, dy = f(op.dx), f(op.dy)
        if not is_zero(dx):
            out.append(f"x = x + {env.val(dx)};")
        if not is_zero(dy):
            out.append(f"y = y + {env.val(dy)};")
        _add_grad(env, g, op.dx, "(-lx)", out)
        _add_grad(env, g, op.dy, "(-ly)", out)
    elif isinstance(op, ir.Edge):
        h = op.h0 * ir.OODPPO if op.chroma else op.h0
        t1e = h * ir.fn("tan", op.e)
        psi = op.e - h * 2.0 * op.fringe / ir.fn("cos", op.e) * (ir.E(1.0) + ir.fn("sin", op.e) * ir.fn("sin", op.e))
        t2e = h * ir.fn("tan", psi)
        t1, t2 = f(t1e), f(t2e)
        if not is_zero(t1):
            out.append(f"px = fma(-x, {env.val(t1)}, px);")
        if not is_zero(t2):
            out.append(f"py = fma(y, {env.val(t2)}, py);")
        _add_grad(env, g, t1e, [("lpx", "x")], out)
        _add_grad(env, g, t2e, [("-lpy", "y")], out)
        if not is_zero(t1):
            out.append(f"lx = fma({env.val(t1)}, lpx, lx);")
        if not is_zero(t2):
            out.append(f"ly = fma({env.neg(t2)}, lpy, ly);")
    elif isinstance(op, (ir.ThickQuad, ir.ThickBend)):
        emit_thick_reverse(op, env, out, g)
    elif isinstance(op, ir.Wire):
        wx, wy, kk = env.val(op.wx), env.val(op.wy), env.val(op.k)
        if any(ir.Expr.of(e).params() for e in (op.wx, op.wy, op.k)):
            raise NotImplementedError("gradient with respect to wire parameters")
        out.append(f"{{ ufloat a = {wx} - x, b = {wy} - y; ufloat r2 = fma(a, a, b * b); "
                   f"ufloat kr = {kk} / r2; px = fma(-kr, a, px); py = fma(-kr, b, py); "
                   f"ufloat q = kr / r2; ufloat faa = q * fma(b, b, -a * a); "
                   f"ufloat fab = {env.lit(-2.0)} * q * a * b; "
                   f"ufloat nlx = fma(-faa, lpx, fma(-fab, lpy, lx)); "
                   f"ly = fma(-fab, lpx, fma(faa, lpy, ly)); lx = nlx; }}")
    elif isinstance(op, ir.Aperture):
        return
    elif isinstance(op, (ir.Code, ir.Observe)):
        raise NotImplementedError(f"{type(op).__name__} has no adjoint")
    else:
        raise TypeError(f"no adjoint for {op!r}")


def emit_kick_reverse(op: ir.Kick, env: Env, out: List[str], g: _Grad):
    """Inverse kick px += Re w, py -= Im w (the last Horner level fused into
    px/py); lambda_x -= a lpx - b lpy, lambda_y += b lpx + a lpy with
    w' = a + i b; coefficient adjoints
    dL/d kn_n = (-lpx Re z^n + lpy Im z^n) chi / n!,
    dL/d ks_n = (lpx Im z^n + lpy Re z^n) chi / n!  (chi = oodppo or 1)."""
    if codegen.chroma_factored(op, env):
        return emit_kick_reverse_factored(ir.Kick(op.kn, op.ks, False), env, out, g)
    d = kick_coeffs(op, env)
    n = len(d)
    top = -1
    for i, (cr, ci) in enumerate(d):
        if not is_zero(cr) or not is_zero(ci):
            top = i
    if top == 2 and all(is_zero(c) for c in (d[0][0], d[0][1], d[1][0], d[1][1])):
        return _emit_quadratic_reverse(op, d, env, out, g)
    lines: List[str] = []

    def v(e):
        return None if is_zero(e) else env.val(e)

    dr = di = None
    if top >= 0:
        wr, wi = v(d[top][0]), v(d[top][1])
        if top >= 1:
            dr, di = v(env.f(d[top][0] * float(top))), v(env.f(d[top][1] * float(top)))
        for lvl in range(top - 1, -1, -1):
            if lvl >= 1:
                er, ei = v(env.f(d[lvl][0] * float(lvl))), v(env.f(d[lvl][1] * float(lvl)))
                nr = codegen._chain([(dr, "x", False), (di, "y", True)], er)
                ni = codegen._chain([(dr, "y", False), (di, "x", False)], ei)
                a, b = env.new_tmp(), env.new_tmp()
                lines.append(f"ufloat {a} = {nr or env.lit(0.0)}; ufloat {b} = {ni or env.lit(0.0)};")
                dr, di = a, b
            if lvl == 0:
                # px += Re(d0 + w z), py -= Im(d0 + w z), accumulated in place
                c0r, c0i = v(d[0][0]), v(d[0][1])
                px, py = "px", "py"
                if wr is not None:
                    px = _fma(wr, "x", px)
                    py = _fma_neg(wr, "y", py)
                if wi is not None:
                    px = _fma_neg(wi, "y", px)
                    py = _fma_neg(wi, "x", py)
                if c0r is not None:
                    px = f"({px} + {c0r})"
                if c0i is not None:
                    py = f"({py} - {c0i})"
                lines.append(f"px = {px}; py = {py};")
            else:
                cr, ci = v(d[lvl][0]), v(d[lvl][1])
                nr = codegen._chain([(wr, "x", False), (wi, "y", True)], cr)
                ni = codegen._chain([(wr, "y", False), (wi, "x", False)], ci)
                a, b = env.new_tmp(), env.new_tmp()
                lines.append(f"ufloat {a} = {nr or env.lit(0.0)}; ufloat {b} = {ni or env.lit(0.0)};")
                wr, wi = a, b
        if top == 0:
            c0r, c0i = v(d[0][0]), v(d[0][1])
            if c0r is not None:
                lines.append(f"px = px + {c0r};")
            if c0i is not None:
                lines.append(f"py = py - {c0i};")
    # coefficient adjoints with the powers of z (z^0 = 1, z^1 = x + i y)
    if any(ir.Expr.of(c).params() for c in list(op.kn) + list(op.ks)):
        chi = ir.OODPPO if op.chroma else ir.E(1.0)
        zr = zi = None
        for m in range(n):
            if m == 1:
                zr, zi = "x", "y"
            elif m >= 2:
                a, b = env.new_tmp(), env.new_tmp()
                lines.append(f"ufloat {a} = fma({zr}, x, -({zi} * y)); ufloat {b} = fma({zr}, y, {zi} * x);")
                zr, zi = a, b
            scale = chi / float(math.factorial(m))
            if m < len(op.kn) and ir.Expr.of(op.kn[m]).params():
                adj = [("-lpx", env.lit(1.0))] if m == 0 else [("-lpx", zr), ("lpy", zi)]
                _add_grad(env, g, op.kn[m] * scale, adj, lines)
            if m < len(op.ks) and ir.Expr.of(op.ks[m]).params():
                adj = [("lpy", env.lit(1.0))] if m == 0 else [("lpx", zi), ("lpy", zr)]
                _add_grad(env, g, op.ks[m] * scale, adj, lines)
    if top >= 1:
        lx, ly = "lx", "ly"
        if dr is not None:
            lx = _fma_neg(dr, "lpx", lx)
            ly = _fma(dr, "lpy", ly)
        if di is not None:
            lx = _fma(di, "lpy", lx)
            ly = _fma(di, "lpx", ly)
        lines.append(f"lx = {lx}; ly = {ly};")
    if lines:
        out.append("{ " + " ".join(lines) + " }")


def _emit_quadratic_reverse(op: ir.Kick, d, env: Env, out: List[str], g: _Grad):
    """Reverse of a pure second-order kick (a sextupole slice, w = d2 z^2):
    z^2 is formed once and serves both the inverse kick (px += Re d2 z^2,
    py -= Im d2 z^2) and the coefficient adjoint; lambda <- J^T lambda with
    w' = 2 d2 z.  15 operations for a slice with both normal and skew parts
    instead of the 18 of the Horner form; 14 for a normal sextupole slice
    (``_emit_normal_quadratic_reverse``)."""
    d2r = None if is_zero(d[2][0]) else env.val(d[2][0])
    d2i = None if is_zero(d[2][1]) else env.val(d[2][1])
    if d2r is not None and d2i is None and not (len(op.ks) > 2 and ir.Expr.of(op.ks[2]).params()):
        return _emit_normal_quadratic_reverse(op, d2r, env.val(env.f(d[2][0] * 2.0)), env, out, g)
    zr, zi = env.new_tmp(), env.new_tmp()
    lines = [f"ufloat {zr} = fma(x, x, -(y * y)); ufloat {zi} = fma(x, y, y * x);"]
    px, py = "px", "py"
    if d2r is not None:
        px = _fma(d2r, zr, px)
        py = _fma_neg(d2r, zi, py)
    if d2i is not None:
        px = _fma_neg(d2i, zi, px)
        py = _fma_neg(d2i, zr, py)
    lines.append(f"px = {px}; py = {py};")
    if any(ir.Expr.of(c).params() for c in list(op.kn) + list(op.ks)):
        chi = ir.OODPPO if op.chroma else ir.E(1.0)
        scale = chi / 2.0
        if len(op.kn) > 2 and ir.Expr.of(op.kn[2]).params():
            _add_grad(env, g, op.kn[2] * scale, [("-lpx", zr), ("lpy", zi)], lines)
        if len(op.ks) > 2 and ir.Expr.of(op.ks[2]).params():
            _add_grad(env, g, op.ks[2] * scale, [("lpx", zi), ("lpy", zr)], lines)
    dr = None if is_zero(d[2][0]) else env.val(env.f(d[2][0] * 2.0))
    di = None if is_zero(d[2][1]) else env.val(env.f(d[2][1] * 2.0))
    a, b = env.new_tmp(), env.new_tmp()
    nr = codegen._chain([(dr, "x", False), (di, "y", True)], None)
    ni = codegen._chain([(dr, "y", False), (di, "x", False)], None)
    lines.append(f"ufloat {a} = {nr}; ufloat {b} = {ni};")
    lines.append(f"lx = {_fma(b, 'lpy', _fma_neg(a, 'lpx', 'lx'))}; ly = {_fma(b, 'lpx', _fma(a, 'lpy', 'ly'))};")
    out.append("{ " + " ".join(lines) + " }")


def _emit_normal_quadratic_reverse(op: ir.Kick, d2: str, d2x2: str, env: Env, out: List[str], g: _Grad):
    """Reverse of a normal sextupole slice, w = d2 z^2 with d2 real.

    Im z^2 = 2xy: the forward's fma(x, y, y*x) rounds to exactly twice the
    rounded product q = x*y (x*y + round(x*y) lies within a quarter ulp of
    2 round(x*y)), so the inverse kick py -= d2 Im z^2 is fma(2 d2, -q, py),
    bit for bit.  With u = y lpy - x lpx and v = y lpx + x lpy the lambda
    update is lx += 2 d2 u, ly += 2 d2 v, and the coefficient adjoint
    -lpx Re z^2 + lpy Im z^2 equals x u + y v: 14 operations."""
    zr, q, u, v = env.new_tmp(), env.new_tmp(), env.new_tmp(), env.new_tmp()
    lines = [f"ufloat {zr} = fma(x, x, -(y * y)); ufloat {q} = x * y;",
             f"px = {_fma(d2, zr, 'px')}; py = {_fma_neg(d2x2, q, 'py')};",
             f"ufloat {u} = fma(y, lpy, -(x * lpx)); ufloat {v} = fma(y, lpx, x * lpy);"]
    if len(op.kn) > 2 and ir.Expr.of(op.kn[2]).params():
        chi = ir.OODPPO if op.chroma else ir.E(1.0)
        _add_grad(env, g, op.kn[2] * (chi / 2.0), [("x", u), ("y", v)], lines)
    lines.append(f"lx = {_fma(d2x2, u, 'lx')}; ly = {_fma(d2x2, v, 'ly')};")
    out.append("{ " + " ".join(lines) + " }")


def emit_kick_reverse_factored(op: ir.Kick, env: Env, out: List[str], g: _Grad):
    """Reverse of a kick whose 1/(1+dp) factor is applied to the whole sum
    (coefficient-table mode): px += oodppo Re w, py -= oodppo Im w, and the
    lambda update and coefficient adjoints use oodppo * (lpx, lpy)."""
    lines: List[str] = []
    wr, wi = codegen.kick_sum(op, env, lines)
    if wr is not None:
        lines.append(f"px = fma(oodppo, {wr}, px);")
    if wi is not None:
        lines.append(f"py = fma(-oodppo, {wi}, py);")
    lines.append("ufloat slx = oodppo * lpx, sly = oodppo * lpy;")
    if any(ir.Expr.of(c).params() for c in list(op.kn) + list(op.ks)):
        zr = zi = None
        for m in range(max(len(op.kn), len(op.ks))):
            if m == 1:
                zr, zi = "x", "y"
            elif m >= 2:
                a, b = env.new_tmp(), env.new_tmp()
                lines.append(f"ufloat {a} = fma({zr}, x, -({zi} * y)); ufloat {b} = fma({zr}, y, {zi} * x);")
                zr, zi = a, b
            scale = 1.0 / float(math.factorial(m))
            if m < len(op.kn) and ir.Expr.of(op.kn[m]).params():
                adj = [("-slx", env.lit(1.0))] if m == 0 else [("-slx", zr), ("sly", zi)]
                _add_grad(env, g, op.kn[m] * scale, adj, lines)
            if m < len(op.ks) and ir.Expr.of(op.ks[m]).params():
                adj = [("sly", env.lit(1.0))] if m == 0 else [("slx", zi), ("sly", zr)]
                _add_grad(env, g, op.ks[m] * scale, adj, lines)
    dr, di = codegen.kick_derivative(op, env, lines)
    lx, ly = "lx", "ly"
    if dr is not None:
        lx = _fma_neg(dr, "slx", lx)
        ly = _fma(dr, "sly", ly)
    if di is not None:
        lx = _fma(di, "sly", lx)
        ly = _fma(di, "slx", ly)
    if (lx, ly) != ("lx", "ly"):
        lines.append(f"lx = {lx}; ly = {ly};")
    out.append("{ " + " ".join(lines) + " }")


def emit_thick_reverse(op, env: Env, out: List[str], g: _Grad):
    """x_post = M x_pre + d per plane (det M = 1): x_pre = M^-1 (x_post - d),
    the strengths' gradient lambda_post^T (dM x_pre + dd), then
    lambda_pre = M^T lambda_post."""
    if isinstance(op, ir.ThickQuad):
        mx, my = codegen.thick_quad_planes(op)
        dx = dpx = None
    else:
        mx, my = codegen.thick_bend_planes(op)
        dx = None if is_zero(env.f(mx[4])) else mx[4]
        dpx = None if is_zero(env.f(mx[5])) else mx[5]
    for (u, pu, lu, lpu), m, du, dpu in ((("x", "px", "lx", "lpx"), mx, dx, dpx),
                                         (("y", "py", "ly", "lpy"), my, None, None)):
        v11, v12, v21, v22 = (env.val(e) for e in m[:4])
        uu = f"({u} - {env.val(du)})" if du is not None else u
        pp = f"({pu} - {env.val(dpu)})" if dpu is not None else pu
        out.append(f"{{ ufloat a = {uu}, b = {pp}; {u} = fma({v22}, a, -({v12}) * b); "
                   f"{pu} = fma({v11}, b, -({v21}) * a); }}")
        # dL/dtheta += lambda_post^T (dM/dtheta x_pre + dd/dtheta)
        _add_grad(env, g, m[0], [(lu, u)], out)
        _add_grad(env, g, m[1], [(lu, pu)], out)
        _add_grad(env, g, m[2], [(lpu, u)], out)
        _add_grad(env, g, m[3], [(lpu, pu)], out)
        if du is not None:
            _add_grad(env, g, du, lu, out)
        if dpu is not None:
            _add_grad(env, g, dpu, lpu, out)
        out.append(f"{{ ufloat la = {lu}; {lu} = fma({v11}, la, {v21} * {lpu}); "
                   f"{lpu} = fma({v12}, la, {v22} * {lpu}); }}")


# ---------------------------------------------------------------------------
# Kernels
# ---------------------------------------------------------------------------

def _turn_ops(line, flags, parameters):
    return lower(line, flags, parameters)


def merge_drifts(ops, f):
    """Consecutive paraxial drifts as one drift of the summed length, for
    the reverse pass only (``f`` folds an expression): a drift leaves px,
    py, lambda_x and lambda_y unchanged, so the inverse x -= (L1 + L2) px,
    the adjoint lambda_px += (L1 + L2) lambda_x and the length gradient
    lambda_x px (dL1 + dL2) are those of the two drifts up to rounding (the
    reverse reconstructs states by inverse ops, not bit for bit anyway).
    Misalignment shifts and weak-focusing terms that fold to zero (a
    magnet without offsets, a straight magnet) do nothing and are dropped
    first.  The forward kernels keep every op of the reference."""
    out = []
    for op in ops:
        if isinstance(op, ir.Shift) and is_zero(f(op.dx)) and is_zero(f(op.dy)):
            continue
        if isinstance(op, ir.Weak) and is_zero(f(op.w)):
            continue
        if isinstance(op, ir.Drift) and out and isinstance(out[-1], ir.Drift):
            out[-1] = ir.Drift(ir.Expr.of(out[-1].L) + op.L)
        else:
            out.append(op)
    return out


def _chi_factor(e: ir.Expr):
    """A folded derivative as ('const', c) for a constant c, ('chi', c) for
    c * oodppo, or None for anything else (per-particle, table-dependent)."""
    t = e.t
    if t[0] == "c":
        return ("const", t[1])
    if t == ("v", "oodppo"):
        return ("chi", 1.0)
    if t[0] == "neg":
        a = _chi_factor(t[1])
        return None if a is None else (a[0], -a[1])
    if t[0] in ("*", "/"):
        a, b = _chi_factor(t[1]), _chi_factor(t[2])
        if a is None or b is None:
            return None
        if t[0] == "*":
            if a[0] == "chi" and b[0] == "chi":
                return None
            return ("chi" if "chi" in (a[0], b[0]) else "const", a[1] * b[1])
        if b[0] != "const" or b[1] == 0.0:
            return None
        return (a[0], a[1] / b[1])
    return None


def _rotation(ops, head, p, f):
    """Number of leading ops of the period (parameter-free drifts and the
    no-op shifts and weak terms that ``merge_drifts`` drops) to move to the
    end of the loop body, so that the drift ending one repetition and the
    drift starting the next fall in one segment and merge; 0 when that
    merges nothing."""
    period = ops[head:head + p]
    k, drift = 0, False
    for op in period:
        if isinstance(op, ir.Drift) and not ir.Expr.of(op.L).params():
            drift = True
        elif not ((isinstance(op, ir.Shift) and is_zero(f(op.dx)) and is_zero(f(op.dy)))
                  or (isinstance(op, ir.Weak) and is_zero(f(op.w)))):
            break
        k += 1
    if not drift or k >= p:
        return 0
    rest = merge_drifts(period[k:], f)
    return k if rest and isinstance(rest[-1], ir.Drift) else 0


def _scalable(ops, f) -> bool:
    """True when the turn holds only paraxial drifts, kicks scaled by
    1/(1+dp) and apertures once no-op shifts and weak terms are dropped: the
    ops that the reverse pass can run on scaled momenta."""
    for op in merge_drifts(ops, f):
        if isinstance(op, (ir.Drift, ir.Aperture)) or (isinstance(op, ir.Kick) and op.chroma):
            continue
        return False
    return True


def _reverse_layout(ops, env: Env, ntheta: int):
    """Reverse pass of one turn: the ops backwards, the repeated period as a
    descending loop; gradient registers flushed to LDS per segment.

    The loop body is the period rotated past its leading drift, so the
    drift closing one repetition and the one opening the next are a single
    drift of the reverse pass (the last repetition's body and the first's
    leading drift run outside the loop).

    Gradient scaling: the flush of a sum multiplies it by the derivative of
    its coefficient.  When every derivative is a constant times 1/(1+dp)
    (6D kicks on a turn that does not run on scaled momenta), lambda is
    carried multiplied by oodppo instead (the recursion lambda <- J^T lambda
    is linear, so every sum carries the factor) and divided by it again for
    dL/dx0.  When every derivative is then a constant, and one constant
    serves all sums of a parameter, that constant is applied to the
    parameter's total after the reduction (``gscale``) and the flush adds
    the bare sum; otherwise the flush multiplies by the constants.

    Scaled momenta: when the turn is drifts, kicks scaled by 1/(1+dp) and
    apertures (``_scalable``), the reverse pass carries p' = p (1 + dp) and
    lambda_p' = lambda_p / (1 + dp) instead of p and lambda_p.  The kicks
    then use the table's strengths directly (no per-magnet product with
    1/(1+dp) in the loop), each drift's L/(1+dp) is a per-particle
    invariant, and the strength derivatives are constants.  Returns (lines,
    period, {"lam_scaled": bool, "scaled_momenta": bool, "gscale": [ntheta]})."""
    head, p, reps, stride = codegen.find_period(ops, True) if len(ops) > 8 else (0, 0, 0, 0)
    rot = _rotation(ops, head, p, env.f) if reps >= 3 else 0
    lo = head + rot                        # first op of the loop body
    n = reps - 1 if rot else reps          # repetitions inside the loop
    flushes = []     # (placeholder, parameters it adds to, index text, [(sum, plain text, factor)])
    env.scaled_momenta = _scalable(ops, env.f)

    def segment(seg_ops, lv=None):
        g = _Grad()
        body: List[str] = []
        for op in reversed(merge_drifts(seg_ops, env.f)):
            emit_reverse(op, env, body, g)
        decl = [f"ufloat {r} = {env.lit(0.0)};" for r in g.sums.values()]
        flush = []
        for i in g.params():
            terms = [(r, env.val(dd), _chi_factor(env.f(dd))) for (ii, dd), r in g.sums.items() if ii == i]
            j = i - env.shared_offset
            idx = f"({j} + {lv} * {stride})" if (lv and stride) else f"{j}"
            tag = f"@FLUSH{len(flushes)}@"
            covered = sorted({j + r * int(stride) for r in range(n)}) if lv else [j]
            flushes.append((tag, covered, idx, terms))
            flush.append(tag)
        return ["{"] + ["    " + s for s in decl + body + flush] + ["}"]

    out: List[str] = []
    if reps >= 2:
        out += segment(ops[lo + p * n:])
        env.stride = ("rep", stride)
        try:
            inner = segment(ops[lo:lo + p], "rep")
        finally:
            env.stride = None
        pre, top, end = codegen.prefetch_lines(env, str(n - 1), "(rep > 0 ? rep - 1 : 0)")
        out += pre
        out.append(codegen.loop_pragma(env))
        out.append(f"for (int rep = {n - 1}; rep >= 0; --rep) {{")
        out += ["    " + s for s in top + inner + end]
        out.append("}")
        out += segment(ops[:lo])
    else:
        out += segment(ops)

    kinds = {t[2][0] if t[2] else None for _, _, _, terms in flushes for t in terms}
    factored = kinds in (set(), {"const"}, {"chi"})
    gscale = [1.0] * ntheta
    bare = set()      # flushes that add the bare sum
    if factored:
        # a parameter is scaled after the reduction when every flush adding
        # to it is a single sum with one common factor, and every other
        # parameter those flushes add to is scaled likewise
        per = {}
        for k, (_, cov, _, terms) in enumerate(flushes):
            for j in cov:
                per.setdefault(j, []).append(k)
        ok = {j for j, ks in per.items()
              if all(len(flushes[k][3]) == 1 for k in ks) and len({flushes[k][3][0][2][1] for k in ks}) == 1}
        changed = True
        while changed:
            changed = False
            for j in list(ok):
                if any(c not in ok for k in per[j] for c in flushes[k][1]):
                    ok.discard(j)
                    changed = True
        for j in ok:
            gscale[j] = flushes[per[j][0]][3][0][2][1]
        bare = {k for k, (_, cov, _, terms) in enumerate(flushes)
                if all(j in ok for j in cov) or (len(terms) == 1 and terms[0][2][1] == 1.0)}
    text = {}
    for k, (tag, cov, idx, terms) in enumerate(flushes):
        if k in bare:
            acc_ = terms[0][0]
        elif factored:
            acc_ = f"{terms[0][0]} * {env.lit(terms[0][2][1])}"
            for r, _, fc in terms[1:]:
                acc_ = f"fma({r}, {env.lit(fc[1])}, {acc_})"
        else:
            acc_ = f"{terms[0][0]} * {terms[0][1]}"
            for r, dv, _ in terms[1:]:
                acc_ = f"fma({r}, {dv}, {acc_})"
        text[tag] = f"unsafeAtomicAdd(&acc[{idx} * {ACCW} + slot], {acc_});"
    out = [_fill_flush(s, text) for s in out]
    info = {"lam_scaled": factored and kinds == {"chi"}, "scaled_momenta": env.scaled_momenta, "gscale": gscale}
    env.scaled_momenta = False
    return out, (head, p, reps, stride), info


def _fill_flush(line: str, text: dict) -> str:
    if "@FLUSH" not in line:
        return line
    key = line.strip()
    return line.replace(key, text[key])


def _table_kernel(env: Env, flat, period, reps, stride, F, S, name="ufo_adj_table"):
    """Kernel evaluating every coefficient of the table from theta, in the
    arithmetic order the tracking kernels would use inline: lane r < reps
    writes repetition r's period slots, lane 0 also the flat ones."""
    p = Env(fp32=False, consts=dict(env.consts), shared_offset=env.shared_offset)
    lines = []
    if flat:
        lines.append("if (r == 0) {")
        lines += [f"    ct[{k}] = {p.raw(e)};" for k, e in enumerate(flat)]
        lines.append("}")
    if period and reps:
        p.stride = ("r", stride)
        lines.append(f"if (r < {reps}) {{")
        lines += [f"    ct[{F} + {j} + r * {S}] = {p.raw(e)};" for j, e in enumerate(period)]
        lines.append("}")
    return f"""
extern "C" __global__ void __launch_bounds__(64)
{name}(const ufloat *__restrict__ th, ufloat *__restrict__ ct)
{{
    const int r = blockIdx.x * 64 + threadIdx.x;
{chr(10).join('    ' + x for x in lines)}
}}
"""


def adjoint_kernels(line, flags, turns, coords, lattice_parameters, dp=0.0, every=32, block=256):
    """(forward kernel, reverse kernel, table kernel, table layout, number of
    checkpoints) of a loss on the final state.  ``coords`` are the
    per-particle coordinate columns (names among x, px, y, py, dp),
    ``lattice_parameters`` the shared (label-or-index, name) parameters, in
    the order of theta.

    Every coefficient that depends on theta alone is evaluated once per call
    by the table kernel (``ufo_adj_table``) into a table ct[] in global
    memory.  Both tracking kernels index it with wave-uniform offsets only,
    so its values arrive by scalar loads into SGPRs (no LDS instruction in
    the turn loop); inside the period loop the coefficients of the next
    repetition are loaded one iteration ahead."""
    if not flags & DOUBLE_PRECISION:
        raise ValueError("the adjoint tracker computes in FP64 (DOUBLE_PRECISION)")
    coords = list(coords)
    shared = list(lattice_parameters)
    parameters = coords + shared
    ops = _turn_ops(line, flags, parameters)
    for op in ops:
        if isinstance(op, ir.Code) or any(e.uses_var("turn") or e.has_raw() for e in op.exprs()):
            raise NotImplementedError("raw user code has no adjoint")
    nck = (turns + every - 1) // every + 1
    src = codegen._coord_sources(coords, flags, dp)
    table: list = []
    slots: dict = {}

    hd, pp, rp = codegen.find_period(ops) if len(ops) > 8 else (0, 0, 0)
    unroll = codegen.unroll_for(ops[hd:hd + pp]) if rp >= 2 else 1

    def make_env():
        env = codegen._env_for(flags, src, None)
        env.shared_offset = len(coords)
        env.table, env.period_slots = table, slots      # one table for both kernels
        env.table_name = "ct"
        env.prefetch = True
        env.factor_chroma = False
        env.unroll = unroll
        return env

    P = max(1, len(coords))
    nth = max(1, len(shared))
    pre = codegen._PRELUDE.format(ufloat="double", nan='__builtin_nan("")')

    envf = make_env()
    fbody, period = codegen._layout(ops, envf)
    envb = make_env()
    rbody, rperiod, rinfo = _reverse_layout(ops, envb, nth)
    head, p, reps, stride = rperiod
    F, S = len(table), len(slots)
    reps = int(reps) if reps >= 2 and S else 0
    ns = max(1, F + reps * S)
    layout = {"flat": [e for e, _ in table], "period": sorted(slots, key=slots.get), "reps": reps, "gscale": rinfo["gscale"][:len(shared)],
              "stride": int(stride), "F": F, "S": S, "nslots": ns}

    def fill(text):
        return text.replace("@PB@", str(F)).replace("@PS@", str(max(S, 1)))

    load = codegen._load_block(envf, src, P, "row", "    ")
    store_ck = ("{ ufloat *c = ckpt + (size_t)ck * 4 * np + idx; c[0] = x; c[np] = px; c[2 * np] = y; "
                "c[3 * np] = py; }")
    fwd = pre + fill(f"""
extern "C" __global__ void __launch_bounds__(256)
ufo_adj_fwd(const ufloat *__restrict__ pool, const ufloat *__restrict__ ct, ufloat *__restrict__ ckpt,
            ufloat *__restrict__ xout, int *__restrict__ lost, int particles)
{{
    {codegen._decls(envf)}
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= particles) return;
    const size_t np = (size_t)particles;
    const ufloat *__restrict__ row = pool + (size_t)idx * {P};
{chr(10).join(load)}
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
""")
    fk = codegen.Kernel(name="ufo_adj_fwd", src=fwd, kind="adj_fwd", fp32=False,
                        flops_per_turn=codegen._count(ops, envf), elements=codegen.element_count(line), block=256,
                        period=period)

    load = codegen._load_block(envb, src, P, "row", "        ")
    # lambda carried times oodppo when every gradient factor holds it
    ls, us = (" * oodppo", " / oodppo") if rinfo["lam_scaled"] else ("", "")
    # scaled momenta: p' = p (1 + dp), lambda_p' = lambda_p / (1 + dp)
    sm = rinfo["scaled_momenta"]
    lp, up = (ls + " * oodppo", us + " / oodppo") if sm else (ls, us)
    pin = " / oodppo" if sm else ""
    bwd = pre + fill(f"""
extern "C" __global__ void __launch_bounds__({block})
ufo_adj_bwd(const ufloat *__restrict__ pool, const ufloat *__restrict__ ct, const ufloat *__restrict__ ckpt,
            const ufloat *__restrict__ lam, const int *__restrict__ lost, ufloat *__restrict__ gx0,
            ufloat *__restrict__ gpart, int particles)
{{
    __shared__ ufloat acc[{nth} * {ACCW}];
    for (int i = threadIdx.x; i < {nth} * {ACCW}; i += {block}) acc[i] = 0.0;
    __syncthreads();   // every slot cleared before any wave adds to it
    const int idx = blockIdx.x * {block} + threadIdx.x;
    const int slot = threadIdx.x % {ACCW};
    if (idx < particles && lost[idx] == {turns}) {{
        {codegen._decls(envb)}
        const size_t np = (size_t)particles;
        const ufloat *__restrict__ row = pool + (size_t)idx * {P};
{chr(10).join(load)}
        ufloat lx = lam[(size_t)idx * 4]{ls}, lpx = lam[(size_t)idx * 4 + 1]{lp};
        ufloat ly = lam[(size_t)idx * 4 + 2]{ls}, lpy = lam[(size_t)idx * 4 + 3]{lp};
        for (int ck = {nck - 1}; ck >= 1; --ck) {{
            const int t_end = (ck * {every} < {turns}) ? ck * {every} : {turns};
            const int t_begin = (ck - 1) * {every};
            const ufloat *c = ckpt + (size_t)ck * 4 * np + idx;
            x = c[0]; px = c[np]{pin}; y = c[2 * np]; py = c[3 * np]{pin};
            for (int turn = t_end - 1; turn >= t_begin; --turn) {{
{chr(10).join('                ' + s_ for s_ in rbody)}
            }}
        }}
        ufloat *o = gx0 + (size_t)idx * 4;
        o[0] = lx{us}; o[1] = lpx{up}; o[2] = ly{us}; o[3] = lpy{up};
    }}
    __syncthreads();
    for (int j = threadIdx.x; j < {nth}; j += {block}) {{
        ufloat s = 0.0;
        for (int k = 0; k < {ACCW}; ++k) s += acc[j * {ACCW} + k];
        gpart[(size_t)blockIdx.x * {nth} + j] = s;
    }}
}}
""")
    bk = codegen.Kernel(name="ufo_adj_bwd", src=bwd, kind="adj_bwd", fp32=False,
                        flops_per_turn=codegen._count(ops, envb), elements=codegen.element_count(line), block=block,
                        period=rperiod)
    tsrc = pre + _table_kernel(envf, layout["flat"], layout["period"], reps, stride, F, max(S, 1))
    tk = codegen.Kernel(name="ufo_adj_table", src=tsrc, kind="adj_table", fp32=False, flops_per_turn=0,
                        elements=0, block=64)
    return fk, bk, tk, layout, nck


# ---------------------------------------------------------------------------
# Driver and autograd
# ---------------------------------------------------------------------------


class AdjointTracker(Simulation):
    """Tracks particles through ``line`` for ``turns`` turns with the lattice
    parameters ``lattice_parameters`` (tuples (label-or-index, name), the
    reference's parameter convention, ufo/lattice.py:109-110) taken from a
    shared vector theta, and back-propagates a final-state loss.

    ``coords``: per-particle initial-coordinate columns (names among x, px,
    y, py, dp).  ``every``: turns between checkpoints.

    ``forward(x0, theta)`` -> (final state [N, 4] (NaN rows for lost
    particles), lost turn [N]);  ``backward(lam)`` with lam = dL/d(final
    state) -> (dL/d(x, px, y, py)_0 [N, 4], dL/dtheta [P]).
    """

    def __init__(self, line, flags, turns, coords, lattice_parameters, dp=0.0, every=32, context=None,
                 options=None):
        self._init_common(flags, context, options)
        self.turns = int(turns)
        self.coords = list(coords)
        self.lattice_parameters = list(lattice_parameters)
        self.every = int(every)
        self.kfwd, self.kbwd, self.ktab, self.layout, self.ncheck = adjoint_kernels(
            line, flags, self.turns, self.coords, self.lattice_parameters, dp, self.every)
        self.kernel = self.kfwd
        self._mods = None
        self._saved = None
        # per-parameter factors the reverse kernel leaves to the reduction
        gs = self.layout["gscale"]
        self._gscale = (torch.tensor(gs, dtype=torch.float64, device=self.device)
                        if any(v != 1.0 for v in gs) else None)

    def modules(self):
        if self._mods is None:
            self._mods = (self.ctx.module(self.kfwd.src, self.options), self.ctx.module(self.kbwd.src, self.options),
                          self.ctx.module(self.ktab.src, self.options))
        return self._mods

    def coefficients(self, theta):
        """The coefficient table of theta (device tensor [nslots]), computed
        by the table kernel on the current stream."""
        th = self._theta(theta)
        ct = torch.zeros(self.layout["nslots"], dtype=torch.float64, device=self.device)
        if self.layout["F"] or self.layout["reps"]:
            grid = max(1, (max(self.layout["reps"], 1) + 63) // 64)
            self.modules()[2].launch(self.ktab.name, grid, 64, [th, ct])
        return ct

    def _theta(self, theta):
        th = torch.as_tensor(theta, dtype=torch.float64, device=self.device).contiguous().reshape(-1)
        if th.numel() != len(self.lattice_parameters):
            raise ValueError(f"theta has {th.numel()} values for {len(self.lattice_parameters)} lattice parameters")
        if th.numel() == 0:
            th = torch.zeros(1, dtype=torch.float64, device=self.device)
        return th

    def forward(self, x0, theta):
        """x0 [particles, len(coords)], theta [len(lattice_parameters)].
        The buffers the reverse pass needs are kept for the next
        ``backward()`` call without ``saved``."""
        xf, lost, self._saved = self.forward_saved(x0, theta)
        return xf, lost

    def forward_saved(self, x0, theta):
        """As ``forward``, returning (final state, lost, saved) where
        ``saved`` (pool, coefficient table, checkpoints, lost) belongs to this
        pass only: give it to ``backward(lam, saved)`` to differentiate this
        pass even after later forwards."""
        fwd = self.modules()[0]
        x0 = torch.as_tensor(x0, dtype=torch.float64, device=self.device).contiguous()
        if x0.dim() != 2 or x0.shape[1] != len(self.coords):
            raise ValueError(f"x0 has shape {tuple(x0.shape)}; expected (particles, {len(self.coords)})")
        n = x0.shape[0]
        pool = x0 if x0.shape[1] > 0 else torch.zeros((n, 1), dtype=torch.float64, device=self.device)
        ct = self.coefficients(theta)
        ckpt = torch.empty((self.ncheck, 4, n), dtype=torch.float64, device=self.device)
        xf = torch.empty((n, 4), dtype=torch.float64, device=self.device)
        lost = torch.empty((n,), dtype=torch.int32, device=self.device)
        if n:
            fwd.launch(self.kfwd.name, (n + 255) // 256, 256, [pool, ct, ckpt, xf, lost, n])
        return xf, lost, (pool, ct, ckpt, lost)

    def backward(self, lam, saved=None):
        """lam = dL/d(final x, px, y, py) [particles, 4] for the pass whose
        ``saved`` buffers are given (default: the last ``forward``); returns
        (dL/dx0 [particles, 4], dL/dtheta)."""
        bwd = self.modules()[1]
        if saved is None:
            saved = self._saved
        if saved is None:
            raise RuntimeError("backward() needs a forward() first")
        pool, th, ckpt, lost = saved
        n = pool.shape[0]
        lam = torch.nan_to_num(torch.as_tensor(lam, dtype=torch.float64, device=self.device), nan=0.0).contiguous()
        if tuple(lam.shape) != (n, 4):
            raise ValueError(f"lam has shape {tuple(lam.shape)}; expected ({n}, 4)")
        block = self.kbwd.block
        nb = max(1, (n + block - 1) // block)
        npar = max(1, len(self.lattice_parameters))
        gx0 = torch.zeros((n, 4), dtype=torch.float64, device=self.device)
        gpart = torch.empty((nb, npar), dtype=torch.float64, device=self.device)
        if n:
            bwd.launch(self.kbwd.name, nb, block, [pool, th, ckpt, lam, lost, gx0, gpart, n])
        else:
            gpart.zero_()
        g = gpart.sum(dim=0)[:len(self.lattice_parameters)]
        if self._gscale is not None:
            g = g * self._gscale
        return gx0, g


class _TrackFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, x0, tracker):
        xf, lost, saved = tracker.forward_saved(x0, theta)
        ctx.tracker = tracker
        ctx.saved = saved          # this pass's own buffers, whatever the tracker runs next
        ctx.ncoord = x0.shape[1]
        ctx.coords = tracker.coords
        ctx.mark_non_differentiable(lost)
        return xf, lost

    @staticmethod
    def backward(ctx, grad_xf, grad_lost):
        gx0, gth = ctx.tracker.backward(grad_xf, ctx.saved)
        g0 = torch.zeros((gx0.shape[0], ctx.ncoord), dtype=gx0.dtype, device=gx0.device)
        for j, name in enumerate(ctx.coords):
            if name in ("x", "px", "y", "py"):
                g0[:, j] = gx0[:, ("x", "px", "y", "py").index(name)]
        return gth, g0, None


def track(tracker: AdjointTracker, theta, x0):
    """Differentiable tracking: (final state [N, 4], lost turn [N]).
    Gradients flow to ``theta`` (shared lattice parameters) and to the x,
    px, y, py columns of ``x0`` (the dp column gets none)."""
    return _TrackFunction.apply(theta, x0, tracker)


def amplitude_loss(xf, aperture=0.04, weight=1e-5, sharpness=20.0, onset=0.25):
    """Final-state objective: mean over particles of alive * (x^2 + y^2)
    plus ``weight`` times a soft-survival penalty, per particle
    alive * softplus(s (r^2/a^2 - onset)) / s + (1 - alive), which grows as a
    surviving particle nears the aperture a and counts each lost particle as
    1 (lost particles carry no gradient)."""
    alive = torch.isfinite(xf[:, 0]) & torch.isfinite(xf[:, 2])
    x = torch.where(alive, xf[:, 0], torch.zeros_like(xf[:, 0]))
    y = torch.where(alive, xf[:, 2], torch.zeros_like(xf[:, 2]))
    r2 = x * x + y * y
    soft = torch.nn.functional.softplus(sharpness * (r2 / (aperture * aperture) - onset)) / sharpness
    pen = torch.where(alive, soft, torch.ones_like(soft))
    return torch.mean(r2) + weight * torch.mean(pen)
