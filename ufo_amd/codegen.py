"""HIP source emitter: IR ops -> kernels compiled by hipRTC for gfx950.

Kernel shape (all kinds):
  * one particle per lane, wave64; the particle state x, px, y, py lives in
    FP64 (or FP32) VGPRs, z/dp/oodppo and every per-particle invariant (an
    expression of parameters and oodppo: chromatic kick strengths, thick
    element matrices, edge kicks) are computed once per particle and kept in
    registers;
  * element constants are folded on the host and emitted as literals, so they
    end up as SGPR constants of the ISA;
  * arithmetic is written with explicit ``fma`` and compiled with
    ``-ffp-contract=off``: what is emitted is what runs;
  * a turn whose op list is a repetition of one period (a ring of identical
    cells) is emitted as a loop over the period, keeping the code within the
    instruction cache;
  * HBM traffic only at particle start (one row of the parameter pool) and at
    particle end (the loss turn), or at the observation points of Track.

Kernel kinds:
  ``sa``      StableAperture (ufo/sa.py:91-126).  Persistent lanes with
              refill: a lane whose particle is lost (or survives all turns)
              stores the result and claims the next particle from a
              wave-aggregated atomic counter (the reference's
              ``atomic_inc(pool_idx)`` work pool, ufo/sa.py:122), so lanes do
              not idle behind the longest-living particle of their wave.
  ``track``   Track (ufo/track.py:84-114), one lane per particle, samples
              stored structure-of-arrays [sample][coordinate][particle] so the
              stores of a wave are coalesced.
  ``optics``  Optics (ufo/optics.py:87-115): the three-ray linear-map kernel.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import crmath, ir
from .lattice import lower
from .flags import FIVED, DOUBLE_PRECISION
from .ir import Expr, fold, is_zero


COORDS = ("x", "px", "y", "py", "z", "dp")


# ---------------------------------------------------------------------------
# Expression printing
# ---------------------------------------------------------------------------

def literal(v: float, fp32: bool) -> str:
    if math.isnan(v):
        return "__builtin_nanf(\"\")" if fp32 else "__builtin_nan(\"\")"
    if math.isinf(v):
        s = "__builtin_inff()" if fp32 else "__builtin_inf()"
        return s if v > 0 else f"(-{s})"
    s = repr(float(np.float32(v))) if fp32 else repr(float(v))
    if "e" not in s and "." not in s:
        s += ".0"
    return f"{s}f" if fp32 else s


class NeedsFlat(Exception):
    """A period loop cannot be emitted for this op list; emit it flat."""


@dataclass
class Env:
    """Code-generation context of one kernel."""
    fp32: bool
    consts: Dict[str, float]                 # kernel variables known at generation (dp, oodppo)
    invariants: Dict[Expr, str] = field(default_factory=dict)
    inv_code: List[str] = field(default_factory=list)
    params_used: set = field(default_factory=set)
    tmp: int = 0
    param_mem: bool = False          # parameters read from the pool row at use
    shared_offset: int = 1 << 30     # parameter columns >= this come from the shared vector th[]
    table: Optional[list] = None     # coefficient-table mode: [(expr, in_period)] slots of ct[]
    period_slots: Optional[dict] = None
    last_period: tuple = (0, 0, 0, 0)
    diff_cols: Optional[list] = None  # parameter columns with gradients (None: every shared one)
    no_loop: bool = False            # emit the turn flat (no period loop)
    stride: Optional[tuple] = None   # (loop variable, parameter stride) inside a strided period loop
    table_name: str = "ct"           # array holding the coefficient table
    prefetch: bool = False           # period-loop coefficients in registers K<j>, loaded one repetition ahead
    factor_chroma: bool = True       # table mode: apply 1/(1+dp) to whole kick sums
    products: Dict[tuple, str] = field(default_factory=dict)  # prefetch mode: (slot, variable) -> register
    remap_loops: bool = False        # table mode: loop over repetitions whose parameters map column to column
    unroll: int = 1                  # period-loop unroll factor (1: the loop stays rolled)
    scaled_momenta: bool = False     # reverse pass: px, py carried times (1 + dp), lambda_px, lambda_py over it

    def f(self, e) -> Expr:
        return fold(e, self.fp32, self.consts)

    def lit(self, v: float) -> str:
        return literal(v, self.fp32)

    def raw(self, e: Expr) -> str:
        """C text of an expression, without hoisting."""
        t = e.t
        k = t[0]
        if k == "c":
            return self.lit(t[1])
        if k == "p" and self.stride is not None and isinstance(self.stride[1], list):
            if t[1] >= self.shared_offset:
                # inside a remapped period loop every lattice parameter must
                # come from the coefficient table
                raise NeedsFlat(f"parameter column {t[1]} read directly in a remapped loop")
            self.params_used.add(t[1])
            return f"P{t[1]}"
        if k == "p" and t[1] >= self.shared_offset:
            j = t[1] - self.shared_offset
            if self.stride is not None and self.stride[1] != 0:
                return f"th[{j} + {self.stride[0]} * {self.stride[1]}]"
            return f"th[{j}]"
        if k == "p":
            if self.stride is not None and self.stride[1] != 0:
                return f"row[{t[1]} + {self.stride[0]} * {self.stride[1]}]"
            if self.param_mem:
                return f"row[{t[1]}]"
            self.params_used.add(t[1])
            return f"P{t[1]}"
        if k == "v":
            return t[1]
        if k == "r":
            return f"({t[1]})"
        if k == "neg":
            return f"(-{self.raw(t[1])})"
        if k == "f":
            return f"{t[1]}({self.raw(t[2])})"
        if k == "?":
            return f"({self.raw(t[1])} ? {self.raw(t[2])} : {self.raw(t[3])})"
        return f"({self.raw(t[1])} {k} {self.raw(t[2])})"

    def val(self, e) -> str:
        """C text of a folded expression; non-constant expressions that do
        not depend on the turn are hoisted into per-particle invariants."""
        e = self.f(e)
        k = e.t[0]
        if self.table is not None and k != "c" and e.params() and not (e.uses_var("oodppo") or e.uses_var("dp")) \
                and not e.has_raw() and min(e.params()) >= self.shared_offset:
            return self.table_slot(e)
        if k in ("c", "v"):
            return self.raw(e)
        if k == "p":
            return self.raw(e)
        if e.has_raw() or e.uses_var("turn"):
            return self.raw(e)
        in_loop = self.stride is not None and self.stride[1] != 0
        if self.table is not None and in_loop and e.params() and min(e.params()) >= self.shared_offset:
            # a table coefficient combined with per-particle variables: built
            # from the parts, evaluated where it is used
            return self._compose(e)
        if e.params() and (self.param_mem or in_loop):
            return self.raw(e)
        if k == "neg":
            return f"(-{self.val(e.t[1])})"
        name = self.invariants.get(e)
        if name is None:
            # children first, so every shared sub-expression is computed once
            if k == "f":
                if e.t[1] in crmath.FUNCS:      # once per particle: correctly rounded
                    text = crmath.call(e.t[1], self.val(e.t[2]), self.fp32)
                else:
                    text = f"{e.t[1]}({self.val(e.t[2])})"
            elif k == "?":
                text = f"({self.val(e.t[1])} ? {self.val(e.t[2])} : {self.val(e.t[3])})"
            else:
                text = f"({self.val(e.t[1])} {k} {self.val(e.t[2])})"
            name = f"I{len(self.invariants)}"
            self.invariants[e] = name
            self.inv_code.append(f"{name} = {text};")
        return name

    def _compose(self, e) -> str:
        k = e.t[0]
        # (X * Y) / 2^n and (X * Y) * 2^n, X a table coefficient, Y
        # per-particle: the power-of-two scaling moves onto X (exact), so the
        # table holds it and one multiply remains
        if k in ("/", "*") and e.t[2].is_const() and _pow2(e.t[2].value) and e.t[1].t[0] == "*":
            a, b = e.t[1].t[1], e.t[1].t[2]
            for X, Y in ((a, b), (b, a)):
                if X.params() and not (X.uses_var("oodppo") or X.uses_var("dp")) and not X.has_raw() \
                        and min(X.params()) >= self.shared_offset:
                    return self._compose(ir.Expr(("*", ir.Expr((k, X, e.t[2])), Y)))
        if k == "*" and self.prefetch:
            # table coefficient times a kernel variable: the product itself
            # is the loop-carried register, formed when the coefficient
            # arrives, off the critical path of the next iteration
            a, b = e.t[1], e.t[2]
            for X, Y in ((a, b), (b, a)):
                if Y.t[0] == "v" and X.params() and not X.has_raw() and not (X.uses_var("oodppo") or X.uses_var("dp")) \
                        and min(X.params()) >= self.shared_offset:
                    kreg = self.table_slot(X)
                    if kreg.startswith("K"):
                        key = (int(kreg[1:]), Y.t[1])
                        if key not in self.products:
                            self.products[key] = f"KP{len(self.products)}"
                        return self.products[key]
        if k == "neg":
            return f"(-{self.val(e.t[1])})"
        if k == "f":
            return f"{e.t[1]}({self.val(e.t[2])})"
        if k == "?":
            return f"({self.val(e.t[1])} ? {self.val(e.t[2])} : {self.val(e.t[3])})"
        return f"({self.val(e.t[1])} {k} {self.val(e.t[2])})"

    def table_slot(self, e) -> str:
        """Slot of a coefficient in the table; inside a strided period loop
        the slot repeats per repetition (register K<j> under prefetch)."""
        t = self.table_name
        if self.stride is not None and self.stride[1] != 0:
            slots = self.period_slots
            if e not in slots:
                slots[e] = len(slots)
            if self.prefetch:
                return f"K{slots[e]}"
            return f"{t}[@PB@ + {slots[e]} + {self.stride[0]} * @PS@]"
        for k, (ex, _) in enumerate(self.table):
            if ex == e:
                return f"{t}[{k}]"
        self.table.append((e, None))
        return f"{t}[{len(self.table) - 1}]"

    def neg(self, e) -> str:
        e = self.f(e)
        if e.is_const():
            return self.lit(-e.value)
        return self.val(-e)

    def new_tmp(self) -> str:
        self.tmp += 1
        return f"t{self.tmp}"

    def text_invariant(self, key: str, text: str) -> str:
        """A per-particle invariant given as C text (keyed by ``key``)."""
        k = Expr(("r", f"@inv:{key}"))
        name = self.invariants.get(k)
        if name is None:
            name = f"I{len(self.invariants)}"
            self.invariants[k] = name
            self.inv_code.append(f"{name} = {text};")
        return name


# ---------------------------------------------------------------------------
# Forward pass of each op
# ---------------------------------------------------------------------------

def _fma(a, b, c):
    return f"fma({a}, {b}, {c})"


def emit_op(op, env: Env, out: List[str], observe=None):
    f = env.f
    if isinstance(op, ir.Drift):
        L = f(op.L)
        if is_zero(L):
            return
        v = env.val(L)
        out.append(f"x = {_fma('px', v, 'x')}; y = {_fma('py', v, 'y')};")
    elif isinstance(op, ir.ExactDrift):
        L = env.val(op.L)
        opdp2 = env.val(ir.E(1.0 + ir.DP) * (1.0 + ir.DP))
        t = env.new_tmp()
        out.append(f"{{ ufloat {t} = {L} / sqrt({_fma('-py', 'py', _fma('-px', 'px', opdp2))}); "
                   f"x = {_fma('px', t, 'x')}; y = {_fma('py', t, 'y')}; }}")
    elif isinstance(op, ir.Kick):
        emit_kick(op, env, out)
    elif isinstance(op, ir.Weak):
        w = f(op.w)
        if not is_zero(w):
            out.append(f"px = {_fma(env.neg(w), 'x', 'px')};")
    elif isinstance(op, ir.Shift):
        dx, dy = f(op.dx), f(op.dy)
        if not is_zero(dx):
            out.append(f"x = x - {env.val(dx)};")
        if not is_zero(dy):
            out.append(f"y = y - {env.val(dy)};")
    elif isinstance(op, ir.Edge):
        h = op.h0 * ir.OODPPO if op.chroma else op.h0
        t1 = f(h * ir.fn("tan", op.e))
        psi = op.e - h * 2.0 * op.fringe / ir.fn("cos", op.e) * (ir.E(1.0) + ir.fn("sin", op.e) * ir.fn("sin", op.e))
        t2 = f(h * ir.fn("tan", psi))
        if not is_zero(t1):
            out.append(f"px = {_fma('x', env.val(t1), 'px')};")
        if not is_zero(t2):
            out.append(f"py = {_fma('-y', env.val(t2), 'py')};")
    elif isinstance(op, ir.ThickQuad):
        emit_thick_quad(op, env, out)
    elif isinstance(op, ir.ThickBend):
        emit_thick_bend(op, env, out)
    elif isinstance(op, ir.Wire):
        wx, wy, k = env.val(op.wx), env.val(op.wy), env.val(op.k)
        a, b, c = env.new_tmp(), env.new_tmp(), env.new_tmp()
        out.append(f"{{ ufloat {a} = {wx} - x; ufloat {b} = {wy} - y; "
                   f"ufloat {c} = {k} / {_fma(a, a, f'{b} * {b}')}; "
                   f"px = {_fma(c, a, 'px')}; py = {_fma(c, b, 'py')}; }}")
    elif isinstance(op, ir.Aperture):
        # verbatim splice: the window's precedence against "== 0" is C's
        out.append(f"if ({op.window} == 0) x = UFO_NAN;")
    elif isinstance(op, ir.Code):
        out.append("{ " + op.text + " }")
    elif isinstance(op, ir.Observe):
        if observe is None:
            raise ValueError("observation point in a kernel without dumps")
        out.append(observe)
    else:
        raise TypeError(f"no emitter for {op!r}")


def kick_coeffs(op: ir.Kick, env: Env):
    """Folded Horner coefficients d_n = c_n (times oodppo) / n!."""
    n = max(len(op.kn), len(op.ks))
    out = []
    for i in range(n):
        cr = op.kn[i] if i < len(op.kn) else ir.ZERO
        ci = op.ks[i] if i < len(op.ks) else ir.ZERO
        if op.chroma:
            cr = cr * ir.OODPPO
            ci = ci * ir.OODPPO
        fact = float(math.factorial(i))
        if i > 1:
            cr = cr / fact
            ci = ci / fact
        out.append((env.f(cr), env.f(ci)))
    return out


def chroma_factored(op, env: Env) -> bool:
    """In coefficient-table mode a per-particle 1/(1+dp) cannot be folded
    into the per-seed coefficients: the kick is computed without it and
    scaled as a whole (every coefficient carries the same factor)."""
    return (env.table is not None and env.factor_chroma and isinstance(op, ir.Kick) and op.chroma
            and not env.f(ir.OODPPO).is_const())


def emit_kick(op: ir.Kick, env: Env, out: List[str]):
    """w = sum d_n z^n by Horner, px -= Re w, py += Im w; the last level
    accumulates straight into px/py."""
    if chroma_factored(op, env):
        return emit_kick_factored(ir.Kick(op.kn, op.ks, False), env, out)
    d = kick_coeffs(op, env)
    top = -1
    for i, (cr, ci) in enumerate(d):
        if not is_zero(cr) or not is_zero(ci):
            top = i
    if top < 0:
        return
    if top == 0:
        cr, ci = d[0]
        if not is_zero(cr):
            out.append(f"px = px - {env.val(cr)};")
        if not is_zero(ci):
            out.append(f"py = py + {env.val(ci)};")
        return
    lines = []
    wr = None if is_zero(d[top][0]) else env.val(d[top][0])
    wi = None if is_zero(d[top][1]) else env.val(d[top][1])
    for lvl in range(top - 1, 0, -1):
        dr, di = d[lvl]
        dr = None if is_zero(dr) else env.val(dr)
        di = None if is_zero(di) else env.val(di)
        # new_r = wr*x - wi*y + dr ; new_i = wr*y + wi*x + di
        r = _chain([(wr, "x", False), (wi, "y", True)], dr)
        i = _chain([(wr, "y", False), (wi, "x", False)], di)
        tr, ti = env.new_tmp(), env.new_tmp()
        lines.append(f"ufloat {tr} = {r if r else env.lit(0.0)};")
        lines.append(f"ufloat {ti} = {i if i else env.lit(0.0)};")
        wr, wi = tr, ti
    d0r = None if is_zero(d[0][0]) else env.val(d[0][0])
    d0i = None if is_zero(d[0][1]) else env.val(d[0][1])
    # px -= wr*x - wi*y + d0r ; py += wr*y + wi*x + d0i
    px = "px"
    if wr is not None:
        px = _fma_neg(wr, "x", px)
    if wi is not None:
        px = _fma(wi, "y", px)
    if d0r is not None:
        px = f"({px} - {d0r})"
    py = "py"
    if wr is not None:
        py = _fma(wr, "y", py)
    if wi is not None:
        py = _fma(wi, "x", py)
    if d0i is not None:
        py = f"({py} + {d0i})"
    if px != "px":
        lines.append(f"px = {px};")
    if py != "py":
        lines.append(f"py = {py};")
    out.append("{ " + " ".join(lines) + " }")


def kick_sum(op: ir.Kick, env: Env, lines: List[str]):
    """(Re, Im) C names of w = sum d_n z^n (None for a zero part)."""
    d = kick_coeffs(op, env)
    top = -1
    for i, (cr, ci) in enumerate(d):
        if not is_zero(cr) or not is_zero(ci):
            top = i
    if top < 0:
        return None, None
    wr = None if is_zero(d[top][0]) else env.val(d[top][0])
    wi = None if is_zero(d[top][1]) else env.val(d[top][1])
    for lvl in range(top - 1, -1, -1):
        dr = None if is_zero(d[lvl][0]) else env.val(d[lvl][0])
        di = None if is_zero(d[lvl][1]) else env.val(d[lvl][1])
        r = _chain([(wr, "x", False), (wi, "y", True)], dr)
        i = _chain([(wr, "y", False), (wi, "x", False)], di)
        tr, ti = env.new_tmp(), env.new_tmp()
        lines.append(f"ufloat {tr} = {r if r else env.lit(0.0)};")
        lines.append(f"ufloat {ti} = {i if i else env.lit(0.0)};")
        wr, wi = tr, ti
    return wr, wi


def kick_derivative(op: ir.Kick, env: Env, lines: List[str]):
    """(Re, Im) C names of w'(z) = sum n d_n z^(n-1) (None for zero parts)."""
    d = kick_coeffs(op, env)
    top = -1
    for i, (cr, ci) in enumerate(d):
        if not is_zero(cr) or not is_zero(ci):
            top = i
    if top < 1:
        return None, None
    e = [(env.f(d[n + 1][0] * float(n + 1)), env.f(d[n + 1][1] * float(n + 1))) for n in range(top)]
    wr = None if is_zero(e[-1][0]) else env.val(e[-1][0])
    wi = None if is_zero(e[-1][1]) else env.val(e[-1][1])
    for lvl in range(top - 2, -1, -1):
        er = None if is_zero(e[lvl][0]) else env.val(e[lvl][0])
        ei = None if is_zero(e[lvl][1]) else env.val(e[lvl][1])
        r = _chain([(wr, "x", False), (wi, "y", True)], er)
        i = _chain([(wr, "y", False), (wi, "x", False)], ei)
        tr, ti = env.new_tmp(), env.new_tmp()
        lines.append(f"ufloat {tr} = {r if r else env.lit(0.0)};")
        lines.append(f"ufloat {ti} = {i if i else env.lit(0.0)};")
        wr, wi = tr, ti
    return wr, wi


def emit_kick_factored(op: ir.Kick, env: Env, out: List[str]):
    """px -= oodppo Re w, py += oodppo Im w with the factor outside the sum."""
    lines: List[str] = []
    wr, wi = kick_sum(op, env, lines)
    if wr is not None:
        lines.append(f"px = fma(-oodppo, {wr}, px);")
    if wi is not None:
        lines.append(f"py = fma(oodppo, {wi}, py);")
    if lines:
        out.append("{ " + " ".join(lines) + " }")


def _negs(s: str) -> str:
    return s[1:] if s.startswith("-") else f"-{s}"


def _pow2(v: float) -> bool:
    m, _ = math.frexp(abs(v))
    return v != 0.0 and m == 0.5


def _fma_neg(coef: str, var: str, acc: str) -> str:
    """acc - coef * var with the sign on the variable operand (a free source
    modifier; a negated product would cost the compiler a second multiply).
    fma(c, -v, a) and fma(-c, v, a) round identically."""
    if coef.startswith("-"):
        return _fma(coef[1:], var, acc)
    return _fma(coef, f"-{var}", acc)


def _chain(terms, add):
    """a1*b1 (+/-) a2*b2 + add as nested fma; None when every term is zero."""
    acc = add
    for coef, var, negate in reversed(terms):
        if coef is None:
            continue
        if acc is None:
            acc = f"({_negs(coef) if negate else coef} * {var})"
        else:
            acc = _fma_neg(coef, var, acc) if negate else _fma(coef, var, acc)
    return acc


def split_exact(e: Expr, env: Env):
    """(hi, lo) of a constant coefficient whose last operation is done
    exactly: hi + lo equals the exact product/quotient of the two rounded
    operands, so the per-turn product carries no systematic rounding bias
    (the reference multiplies and divides at every turn,
    ``S * px0 / k2``).  Non-constant or single-leaf coefficients: lo = None."""
    f = env.f(e)
    if not f.is_const() or env.fp32:
        return f, None
    t = Expr.of(e).t
    sign = 1.0
    while t[0] in ("neg", "?"):
        if t[0] == "neg":
            sign = -sign
            t = t[1].t
            continue
        c = env.f(t[1])
        if not c.is_const():
            return f, None
        t = (t[2] if c.value != 0.0 else t[3]).t
    if t[0] in ("*", "/"):
        a, b = env.f(t[1]), env.f(t[2])
        if a.is_const() and b.is_const() and b.value != 0.0 and np.isfinite(a.value) and np.isfinite(b.value):
            import mpmath
            with mpmath.workprec(200):
                x = mpmath.mpf(a.value) * mpmath.mpf(b.value) if t[0] == "*" else \
                    mpmath.mpf(a.value) / mpmath.mpf(b.value)
                hi = float(x)
                lo = float(x - mpmath.mpf(hi))
            if hi == sign * f.value * sign and lo != 0.0:
                return ir.const(sign * hi), ir.const(sign * lo)
    return f, None


def split_runtime(e: Expr, env: Env) -> Optional[str]:
    """Low part of a per-particle coefficient hoisted as an invariant whose
    last operation is a product or quotient of two invariants: the exact
    product error fma(a, b, -ab), or the quotient residual
    fma(-q, b, a) / b, so that hi + lo is the exact (to 2^-106) value, as
    ``split_exact`` gives for constants.  None where there is no such part
    (or in FP32, table or in-loop contexts, where coefficients are not
    invariants)."""
    if env.fp32 or env.table is not None or env.param_mem or (env.stride is not None and env.stride[1] != 0):
        return None
    f = env.f(e)
    if f.is_const() or f.has_raw() or f.uses_var("turn"):
        return None
    t = f.t
    if t[0] == "neg":
        lo = split_runtime(t[1], env)
        return None if lo is None else env.text_invariant(f"neg {lo}", f"(-{lo})")
    if t[0] == "?":
        la, lb = split_runtime(t[2], env), split_runtime(t[3], env)
        if la is None and lb is None:
            return None
        c = env.val(t[1])
        zero = env.lit(0.0)
        return env.text_invariant(f"sel {c} {la} {lb}", f"({c} ? {la or zero} : {lb or zero})")
    if t[0] not in ("*", "/"):
        return None
    hi = env.val(f)
    if not hi.startswith("I"):
        return None
    a, b = env.val(t[1]), env.val(t[2])
    if t[0] == "*":
        return env.text_invariant(f"lo* {hi}", f"fma({a}, {b}, -{hi})")
    return env.text_invariant(f"lo/ {hi}", f"(fma(-{hi}, {b}, {a}) / {b})")


def _thick_plane(env: Env, u: str, pu: str, m11, m12, m21, m22, du=None, dpu=None):
    """u' = m11 u + m12 pu (+ du); pu' = m21 u + m22 pu (+ dpu); coefficients
    carry their exact low parts (constants: ``split_exact``; per-particle
    invariants: ``split_runtime``)."""
    a = env.new_tmp()
    pairs = []
    for m in (m11, m12, m21, m22):
        hi, lo = split_exact(m, env)
        if lo is None:
            name = split_runtime(m, env)
            if name is not None:
                lo = Expr(("r", name))
        pairs.append((hi, lo))
    (h11, l11), (h12, l12), (h21, l21), (h22, l22) = pairs

    def row(hu, lu, hp, lp, add):
        acc = None if add is None else env.val(add)
        for coef, var_ in ((lp, pu), (lu, a)):
            if coef is not None:
                acc = _fma(env.val(coef), var_, acc) if acc is not None else f"({var_} * {env.val(coef)})"
        acc = _fma(pu, env.val(hp), acc) if acc is not None else f"{pu} * {env.val(hp)}"
        return _fma(a, env.val(hu), acc)

    return (f"{{ ufloat {a} = {u}; {u} = {row(h11, l11, h12, l12, du)}; "
            f"{pu} = {row(h21, l21, h22, l22, dpu)}; }}")




def thick_quad_planes(op: ir.ThickQuad):
    """(x-plane, y-plane) matrices (m11, m12, m21, m22) of a thick quad,
    ufo/methods.py:106-132: K = sqrt(|k1 oodppo|); focusing in x for k1 > 0,
    else in y (k1 = 0 gives 0/0 = NaN, as in the reference)."""
    k = op.k1
    kk = k * ir.OODPPO if op.chroma else k
    k2 = ir.fn("sqrt", ir.fn("fabs", kk))
    kl = op.L * k2
    C, S = ir.fn("cos", kl), ir.fn("sin", kl)
    CH, SH = ir.fn("cosh", kl), ir.fn("sinh", kl)
    foc = (C, S / k2, -(S * k2), C)
    dfc = (CH, SH / k2, SH * k2, CH)
    cond = ir.gt(k, ir.ZERO)
    fx = tuple(ir.select(cond, a, b) for a, b in zip(foc, dfc))
    fy = tuple(ir.select(cond, b, a) for a, b in zip(foc, dfc))
    return fx, fy


def emit_thick_quad(op: ir.ThickQuad, env: Env, out: List[str]):
    fx, fy = thick_quad_planes(op)
    out.append(_thick_plane(env, "x", "px", *fx))
    out.append(_thick_plane(env, "y", "py", *fy))


def _bend_plane(env: Env, kexpr: Expr, L: Expr, disp: Optional[Expr]):
    """Matrix (and dispersion) of one bend plane with focusing k, following
    the three-way sign branch of ufo/methods.py:157-193."""
    k2 = ir.fn("sqrt", ir.fn("fabs", kexpr))
    kl = L * k2
    C, S = ir.fn("cos", kl), ir.fn("sin", kl)
    CH, SH = ir.fn("cosh", kl), ir.fn("sinh", kl)
    one, zero = ir.E(1.0), ir.ZERO
    pos = [C, S / k2, -(S * k2), C]
    neg = [CH, SH / k2, SH * k2, CH]
    nul = [one, L, zero, one]
    if disp is not None:
        pos += [disp * (one - C) / ir.fn("fabs", kexpr), disp * S / k2]
        neg += [disp * (CH - one) / ir.fn("fabs", kexpr), disp * SH / k2]
        nul += [zero, zero]
    gt, lt = ir.gt(kexpr, ir.ZERO), ir.lt(kexpr, ir.ZERO)
    return [ir.select(gt, p, ir.select(lt, n, z)) for p, n, z in zip(pos, neg, nul)]


def thick_bend_planes(op: ir.ThickBend):
    """(x-plane matrix + dispersion (6 entries), y-plane matrix) of a thick
    sector bend, ufo/methods.py:142-194: y plane with k = k1 (oodppo), x plane
    with k + h^2 / oodppo and dispersion dp h (1-C)/|k|, dp h S/K."""
    h = op.angle / op.L
    k = op.k1
    if op.chroma:
        h = h * ir.OODPPO
        k = k * ir.OODPPO
        kx = k + h * h / ir.OODPPO
    else:
        kx = k + h * h
    # vertical focusing strength is -k: k > 0 defocuses in y (cosh/sinh),
    # k < 0 focuses (cos/sin), ufo/methods.py:157-168
    my = _bend_plane(None, -k, op.L, None)
    mx = _bend_plane(None, kx, op.L, ir.DP * h)
    return mx, my


def emit_thick_bend(op: ir.ThickBend, env: Env, out: List[str]):
    mx, my = thick_bend_planes(op)
    dx = None if is_zero(env.f(mx[4])) else mx[4]
    dpx = None if is_zero(env.f(mx[5])) else mx[5]
    out.append(_thick_plane(env, "x", "px", mx[0], mx[1], mx[2], mx[3], dx, dpx))
    out.append(_thick_plane(env, "y", "py", *my))


# ---------------------------------------------------------------------------
# Turn structure: period detection
# ---------------------------------------------------------------------------

def _signature(op):
    """An op with its parameter indices abstracted, and those indices in a
    fixed order."""
    found = []

    def leaf(e):
        if e.t[0] == "p":
            found.append(e.t[1])
            return Expr(("p", -1))
        return e

    sig = op.map_exprs(lambda e: ir.map_leaves(e, leaf))
    return sig, tuple(found)


def _param_stride(par, head, p, reps):
    """Common per-repetition stride of the parameter indices of a candidate
    period, or None when the repetitions do not differ by one stride."""
    stride = None
    for r in range(1, reps):
        for j in range(p):
            a, b = par[head + j], par[head + r * p + j]
            if len(a) != len(b):
                return None
            for ia, ib in zip(a, b):
                d = ib - ia
                if d % r:
                    return None
                if stride is None:
                    stride = d // r
                elif d != stride * r:
                    return None
    return 0 if stride is None else stride


def _param_maps(par, head, p, reps, fixed_below):
    """Per-repetition maps of parameter columns (repetition 0 -> r), or None
    when a column would map to two columns or a column below
    ``fixed_below`` (per-particle coordinates) would move."""
    maps = [dict() for _ in range(reps)]
    for r in range(reps):
        m = maps[r]
        for j in range(p):
            a, b = par[head + j], par[head + r * p + j]
            if len(a) != len(b):
                return None
            for ia, ib in zip(a, b):
                if ia < fixed_below and ia != ib:
                    return None
                if m.setdefault(ia, ib) != ib:
                    return None
    return maps


def _z_function(s):
    """z[i] = length of the longest common prefix of s and s[i:] (z[0] = 0),
    in linear time (Gusfield's Z algorithm)."""
    n = len(s)
    z = [0] * n
    lo = hi = 0
    for i in range(1, n):
        k = min(hi - i, z[i - lo]) if i < hi else 0
        while i + k < n and s[k] == s[i + k]:
            k += 1
        z[i] = k
        if i + k > hi:
            lo, hi = i, i + k
    return z


def find_period(ops, with_stride=False, remap_from=None):
    """Largest coverage (head, period, repeats[, stride]) of the op list by a
    repeated block starting within the first few ops.  Repetitions may
    differ in their parameter indices by a constant stride per repetition
    (per-magnet parameters of a ring of identical cells).  (0, 0, 0[, 0])
    when nothing repeats."""
    n = len(ops)
    none = (0, 0, 0, 0) if with_stride else (0, 0, 0)
    if n < 2:
        return none
    ids, seen = {}, {}
    sigs, par = [], []
    for op in ops:
        r = seen.get(op)
        if r is None:    # identical ops (the cells of a ring) are abstracted once
            sg, pr = _signature(op)
            r = seen[op] = (ids.setdefault(sg, len(ids)), pr)
        sigs.append(r[0])
        par.append(r[1])
    cands = []
    for head in range(0, min(9, n - 1)):
        s = sigs[head:]
        z = _z_function(s)
        for p in range(1, len(s) // 2 + 1):
            # s[head + p:] repeats s[head:] for z[p] ops: the block of p ops
            # runs (z[p] + p) // p times
            reps = (z[p] + p) // p
            if reps >= 2:
                cands.append((reps * p, -head, -p, head, p, reps))
    cands.sort(reverse=True)
    best, best_cov = none, 0
    for cov, _, _, head, p, reps in cands:
        if cov <= best_cov:
            break
        # shrink the repeat count until the parameter indices follow a
        # stride (or, with remap_from set, a column-to-column map per
        # repetition; columns below remap_from stay fixed)
        while reps >= 2 and reps * p > best_cov:
            st = _param_stride(par, head, p, reps)
            if st is None and remap_from is not None:
                st = _param_maps(par, head, p, reps, remap_from)
            if st is not None:
                best, best_cov = ((head, p, reps, st) if with_stride else (head, p, reps)), reps * p
                break
            reps -= 1
    return best


# ---------------------------------------------------------------------------
# Kernel assembly
# ---------------------------------------------------------------------------

_PRELUDE = """// generated by ufo_amd.codegen — do not edit
typedef {ufloat} ufloat;
#define UFO_NAN {nan}
#ifndef NAN
#define NAN UFO_NAN
#endif
#ifndef UFO_STREAM_STORE
#define UFO_STREAM_STORE(v, p) __builtin_nontemporal_store((v), (p))
#endif
"""


@dataclass
class Kernel:
    """Generated source plus what the driver needs to launch it."""
    name: str
    src: str
    kind: str
    fp32: bool
    flops_per_turn: int        # canonical FLOPs per particle-turn
    elements: int              # flattened elements excluding markers/apertures
    dumps_per_turn: int = 0
    block: int = 256
    period: tuple = (0, 0, 0)
    turns: int = 0             # turns compiled in (Track: the last turn of a run)

    def __post_init__(self):
        self.src = crmath.attach(self.src)


def _layout(body_ops, env, observe=None, loop_var="rep", emit=None, auto_unroll=False):
    """Emit the op list, as a loop over the repeated period when there is
    one.  ``emit(op, env, out, observe)`` writes one op (default: the
    forward pass).  With ``auto_unroll`` a short period is unrolled
    (``unroll_for``): the loop's own instructions and its branch then come
    once per several repetitions, which matters when a repetition is a few
    dozen FP64 instructions (a FODO cell) and the wave's operations form one
    dependent chain."""
    emit = emit or emit_op
    remap_from = env.shared_offset if (env.remap_loops and env.table is not None) else None
    head, p, reps, stride = (find_period(body_ops, True, remap_from) if len(body_ops) > 8 and not env.no_loop
                             else (0, 0, 0, 0))
    if reps >= 2 and isinstance(stride, list):
        # a remapped loop works only if every lattice parameter of the
        # period reaches the table: try on a copy, fall back to flat
        import copy
        trial = copy.deepcopy(env)
        trial.remap_loops = False
        try:
            trial.stride = ("rep", stride)
            for op in body_ops[head:head + p]:
                emit(op, trial, [], observe)
        except NeedsFlat:
            head, p, reps, stride = 0, 0, 0, 0
    env.last_period = (head, p, reps, stride)
    if auto_unroll and reps >= 2:
        env.unroll = unroll_for(body_ops[head:head + p])
    out: List[str] = []
    if reps >= 2:
        for op in body_ops[:head]:
            emit(op, env, out, observe)
        inner: List[str] = []
        env.stride = (loop_var, stride)
        try:
            for op in body_ops[head:head + p]:
                emit(op, env, inner, observe)
        finally:
            env.stride = None
        nxt = f"({loop_var} + 1 < {reps} ? {loop_var} + 1 : {loop_var})"
        pre, top, end = prefetch_lines(env, "0", nxt)
        out += pre
        out.append(loop_pragma(env))
        out.append(f"for (int {loop_var} = 0; {loop_var} < {reps}; ++{loop_var}) {{")
        out += ["    " + s for s in top + inner + end]
        out.append("}")
        for op in body_ops[head + p * reps:]:
            emit(op, env, out, observe)
    else:
        for op in body_ops:
            emit(op, env, out, observe)
    return out, (head, p, reps)


def loop_pragma(env: Env) -> str:
    """Unrolling of the period loop: off by default (the loop body of a ring
    period already fills the instruction stream); a small body may be
    unrolled ``env.unroll`` times so the scheduler interleaves neighbouring
    repetitions."""
    if env.unroll > 1:
        return f"#pragma clang loop unroll_count({int(env.unroll)})"
    return "#pragma clang loop unroll(disable)"


def unroll_for(period_ops, budget=1024, most=4) -> int:
    """Unroll factor for a period loop: as many repetitions (up to ``most``)
    as fit in ``budget`` canonical FLOPs."""
    f = sum(ir.op_flops(op) for op in period_ops) if period_ops else budget
    return int(max(1, min(most, budget // max(f, 1))))


def prefetch_lines(env: Env, first: str, nxt: str):
    """Register rotation of the period-loop table coefficients (prefetch
    mode): K<j> holds repetition ``first``'s value before the loop; each
    iteration loads the next repetition's N<j> at its top (so the load is in
    flight while the body runs) and moves it into K<j> at its end.  Returns
    (before the loop, loop top, loop end) lines; empty without prefetch.
    The repetition index is wave-uniform, so with the table in global
    memory behind a const __restrict__ pointer the loads are scalar loads
    into SGPRs, which the compiler is left free to schedule."""
    if not env.prefetch or not env.period_slots:
        return [], [], []
    t = env.table_name
    js = sorted(env.period_slots.values())
    pre = [f"ufloat K{j} = {t}[@PB@ + {j} + ({first}) * @PS@];" for j in js]
    top = [f"const int nrep_ = {nxt};"] + [f"const ufloat N{j} = {t}[@PB@ + {j} + nrep_ * @PS@];" for j in js]
    end = [f"K{j} = N{j};" for j in js]
    for (j, var), name in sorted(env.products.items(), key=lambda kv: kv[1]):
        pre.append(f"ufloat {name} = K{j} * {var};")
        end.append(f"{name} = K{j} * {var};")
    return pre, top, end


def _coord_sources(parameters, flags, dp):
    """Initial value of each coordinate (ufo/track.py:93-101): a parameter
    column, or 0; FIVED forces z = 0 and dp = the literal."""
    src = {}
    for c in COORDS:
        src[c] = ("p", list(parameters).index(c)) if c in parameters else ("c", 0.0)
    if flags & FIVED:
        src["z"] = ("c", 0.0)
        src["dp"] = ("c", float(dp))
    return src


PARAM_REGISTERS = 24   # at most this many parameters are kept in registers per particle


def count_params(ops):
    used = set()
    for op in ops:
        for e in op.exprs():
            used |= e.params()
    return used


def _env_for(flags, src, ops=None):
    fp32 = not (flags & DOUBLE_PRECISION)
    consts = {}
    if src["dp"][0] == "c":
        dpv = src["dp"][1]
        consts["dp"] = dpv
        if fp32:
            one = np.float32(1.0)
            consts["oodppo"] = float(one / (np.float32(dpv) + one))
        else:
            consts["oodppo"] = 1.0 / (dpv + 1.0)
    env = Env(fp32=fp32, consts=consts)
    if ops is not None and len(count_params(ops)) > PARAM_REGISTERS:
        env.param_mem = True
    return env


def _load_block(env: Env, src, P: int, row: str, indent: str):
    """Statements that load one particle: coordinates, parameters used by the
    body, then the hoisted invariants (which may use both)."""
    lines = []
    for c in COORDS:
        kind, v = src[c]
        if kind == "p":
            lines.append(f"{c} = {row}[{v}];")
        else:
            lines.append(f"{c} = {env.lit(v)};")
    if "oodppo" in env.consts:
        lines.append(f"oodppo = {env.lit(env.consts['oodppo'])};")
    else:
        one = env.lit(1.0)
        lines.append(f"oodppo = {one} / (dp + {one});")
    for i in sorted(env.params_used):
        if i >= P:
            raise IndexError(f"parameters[{i}] used but only {P} parameter columns exist")
        lines.append(f"P{i} = {row}[{i}];")
    lines += env.inv_code
    return [indent + s for s in lines]


def _decls(env: Env):
    names = ["x", "px", "y", "py", "z", "dp", "oodppo"]
    names += [f"P{i}" for i in sorted(env.params_used)]
    names += list(env.invariants.values())
    return "ufloat " + ", ".join(names) + ";"


def _count(ops, env):
    per_op = {}    # identical ops (the cells of a ring) are folded once
    flops = 0
    for op in ops:
        f = per_op.get(op)
        if f is None:
            f = per_op[op] = ir.op_flops(op.map_exprs(env.f))
        flops += f
    return flops


def element_count(line) -> int:
    from .elements import Marker, Aperture
    return sum(1 for e in line.flatten() if not isinstance(e, (Marker, Aperture)))


def sa_kernel(line, flags, turns, parameters, dp, block=256, name="ufo_sa") -> Kernel:
    """StableAperture: lost[p] = number of turns completed before the
    end-of-turn test |x| > 1 || |y| > 1 || isnan(x) fired, or ``turns``
    (ufo/sa.py:116-121)."""
    ops = lower(line, flags, parameters)
    src = _coord_sources(parameters, flags, dp)
    env = _env_for(flags, src, ops)
    body, period = _layout(ops, env, auto_unroll=True)
    P = max(1, len(parameters))
    one = env.lit(1.0)
    load = _load_block(env, src, P, "row", "            ")
    code = _PRELUDE.format(ufloat="float" if env.fp32 else "double",
                           nan='__builtin_nanf("")' if env.fp32 else '__builtin_nan("")')
    code += f"""
extern "C" __global__ void __launch_bounds__({block})
{name}(const ufloat *__restrict__ pool, unsigned int *__restrict__ lost,
       int *__restrict__ counter, int particles, int first, const int *__restrict__ order)
{{
    // lanes start on claims [0, first); further claims come from counter
    // (zeroed by the host before the launch).  Claim c tracks particle
    // order[c] (a permutation of the particles), or particle c without order.
    {_decls(env)}
    int claim = blockIdx.x * {block} + threadIdx.x;
    int idx = 0;
    int turn = 0;
    bool need = true;
    const ufloat *__restrict__ row = pool;
    for (;;) {{
        if (need) {{
            if (claim >= particles) break;
            idx = order ? order[claim] : claim;
            row = pool + (size_t)idx * {P};
{chr(10).join(load)}
            turn = 0;
            need = false;
        }}
{chr(10).join('        ' + s for s in body)}
        bool gone = fabs(x) > {one} || fabs(y) > {one} || isnan(x);
        bool done = gone || turn + 1 >= {turns};
        if (done) {{
            lost[idx] = gone ? turn : {turns};
            unsigned long long m = __ballot(1);
            int lane = __lane_id();
            int leader = __ffsll((long long)m) - 1;
            int rank = __popcll(m & ((1ull << lane) - 1ull));
            int base = 0;
            if (lane == leader) base = atomicAdd(counter, __popcll(m));
            base = __shfl(base, leader);
            claim = first + base + rank;
            need = true;
        }} else {{
            ++turn;
        }}
    }}
}}
"""
    return Kernel(name=name, src=code, kind="sa", fp32=env.fp32, flops_per_turn=_count(ops, env),
                  elements=element_count(line), block=block, period=period)


def track_kernel(line, flags, turns, parameters, where, dp, block=256, name="ufo_track") -> Kernel:
    """Track: samples at the fracture points of ``where`` (and at the end of
    the line for -1), in lattice order, turn after turn; a particle stops
    after the turn at whose end x is NaN (ufo/track.py:62-114).  Output is
    [sample][coordinate][particle].

    One launch runs turns [turn0, turn1).  A launch that starts after turn 0
    resumes from ``state`` ([4][particles]: x, px, y, py; z and dp never
    change and are loaded from the pool as at turn 0) and tracks nothing for
    a particle already lost; a launch that ends before the last turn saves the
    state.  The launch writes every sample of its turns: those after the
    particle's loss turn are NaN (the reference leaves them unwritten).
    Splitting the turns over launches changes no result, and lets the caller
    copy finished samples back while later turns run."""
    where = list(where) if isinstance(where, (list, tuple)) else [where]
    ops = lower(line, flags, parameters, where)
    if -1 in where:
        ops.append(ir.Observe())
    dumps = sum(1 for op in ops if isinstance(op, ir.Observe))
    src = _coord_sources(parameters, flags, dp)
    env = _env_for(flags, src, ops)
    # samples are written once and read back only by the copy to the host:
    # streaming (non-temporal) stores
    store = ("{ size_t b = (size_t)s * 6 * np + idx; "
             "UFO_STREAM_STORE(x, tracks + b); UFO_STREAM_STORE(px, tracks + b + np); "
             "UFO_STREAM_STORE(y, tracks + b + 2 * np); UFO_STREAM_STORE(py, tracks + b + 3 * np); "
             "UFO_STREAM_STORE(z, tracks + b + 4 * np); UFO_STREAM_STORE(dp, tracks + b + 5 * np); ++s; }")
    body, period = _layout(ops, env, observe=store, auto_unroll=True)
    P = max(1, len(parameters))
    load = _load_block(env, src, P, "row", "    ")
    code = _PRELUDE.format(ufloat="float" if env.fp32 else "double",
                           nan='__builtin_nanf("")' if env.fp32 else '__builtin_nan("")')
    code += f"""
extern "C" __global__ void __launch_bounds__({block})
{name}(const ufloat *__restrict__ pool, ufloat *__restrict__ tracks, ufloat *__restrict__ state,
     int particles, int turn0, int turn1)
{{
    {_decls(env)}
    const int idx = blockIdx.x * {block} + threadIdx.x;
    if (idx >= particles) return;
    const size_t np = (size_t)particles;
    const ufloat *__restrict__ row = pool + (size_t)idx * {P};
{chr(10).join(load)}
    bool alive = true;
    if (turn0 > 0) {{
        x = state[idx]; px = state[np + idx]; y = state[2 * np + idx]; py = state[3 * np + idx];
        alive = !isnan(x);
    }}
    int s = turn0 * {dumps};
    if (alive) {{
        for (int turn = turn0; turn < turn1; ++turn) {{
{chr(10).join('            ' + s_ for s_ in body)}
            if (isnan(x)) break;
        }}
    }}
    // the samples of this launch after the particle's loss turn are NaN
    if (s < turn1 * {dumps}) {{
        const ufloat q = UFO_NAN;
        for (; s < turn1 * {dumps}; ++s) {{
            size_t b = (size_t)s * 6 * np + idx;
            UFO_STREAM_STORE(q, tracks + b); UFO_STREAM_STORE(q, tracks + b + np);
            UFO_STREAM_STORE(q, tracks + b + 2 * np); UFO_STREAM_STORE(q, tracks + b + 3 * np);
            UFO_STREAM_STORE(q, tracks + b + 4 * np); UFO_STREAM_STORE(q, tracks + b + 5 * np);
        }}
    }}
    if (alive && turn1 < {turns}) {{
        state[idx] = x; state[np + idx] = px; state[2 * np + idx] = y; state[3 * np + idx] = py;
    }}
}}
"""
    return Kernel(name=name, src=code, kind="track", fp32=env.fp32, flops_per_turn=_count(ops, env),
                  elements=element_count(line), dumps_per_turn=dumps, block=block, period=period,
                  turns=int(turns))


def optics_kernel(line, flags, where, parameters, name="ufo_optics") -> Kernel:
    """Optics: three rays from pool row 0 — x=y=1, px=py=1, dp=1 (plus a
    'dp' parameter if named) — dumped at the observation points
    (ufo/optics.py:87-115).  Output [3][count][6]; only x, px, y, py are
    written."""
    ops = lower(line, flags, parameters, where)
    if -1 in where:
        ops.append(ir.Observe())
    dumps = sum(1 for op in ops if isinstance(op, ir.Observe))
    fp32 = not (flags & DOUBLE_PRECISION)
    env = Env(fp32=fp32, consts={})
    store = ("{ ufloat *o = tracks + ((size_t)ray * count + s) * 6; "
             "o[0] = x; o[1] = px; o[2] = y; o[3] = py; ++s; }")
    body, _ = _layout(ops, env, observe=store)
    zero, one = env.lit(0.0), env.lit(1.0)
    lines = [f"x = y = z = px = py = dp = {zero};",
             f"if (ray == 0) {{ x = {one}; y = {one}; }}",
             f"if (ray == 1) {{ px = {one}; py = {one}; }}",
             f"if (ray == 2) dp = {one};"]
    if "dp" in parameters:
        lines.append(f"dp += row[{list(parameters).index('dp')}];")
    lines.append(f"oodppo = {one} / (dp + {one});")
    for i in sorted(env.params_used):
        lines.append(f"P{i} = row[{i}];")
    lines += env.inv_code
    code = _PRELUDE.format(ufloat="float" if fp32 else "double",
                           nan='__builtin_nanf("")' if fp32 else '__builtin_nan("")')
    code += f"""
extern "C" __global__ void __launch_bounds__(64)
{name}(const ufloat *__restrict__ pool, ufloat *__restrict__ tracks, int count)
{{
    {_decls(env)}
    const int ray = threadIdx.x;
    if (ray >= 3) return;
    const ufloat *__restrict__ row = pool;
{chr(10).join('    ' + s_ for s_ in lines)}
    int s = 0;
{chr(10).join('    ' + s_ for s_ in body)}
    (void)s;
}}
"""
    return Kernel(name=name, src=code, kind="optics", fp32=fp32, flops_per_turn=_count(ops, env),
                  elements=element_count(line), dumps_per_turn=dumps, block=64)


HIPRTC_OPTIONS = ("-O3 -std=c++17 -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops "
                  "-Wno-unused-variable -Wno-unused-but-set-variable")
