"""Element-op IR: the lowered form of a lattice that every kernel emitter reads.

The reference lowers a lattice straight to OpenCL-C text (ufo/methods.py,
ufo/elements.py ``method()``, ufo/lattice.py:86-119).  Here the same physics is
lowered to a short list of typed ops whose numeric slots are small expression
trees (:class:`Expr`).  Constant sub-trees fold on the host in the kernel's
precision; sub-trees that depend on a per-particle parameter or on ``oodppo``
become per-particle invariants hoisted out of the turn loop by the emitter.

Every op documents the reference template it follows; the arithmetic of each
op is written in ``codegen.py`` (forward), ``tangent.py`` (JVP) and
``adjoint.py`` (VJP) from the same op fields.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, fields
from typing import Tuple

import mpmath
import numpy as np

from .flags import LINEAR, EXACT, ACHROMATIC

# --------------------------------------------------------------------------
# Expressions
# --------------------------------------------------------------------------

_PARAM_RE = re.compile(r"^\s*\(?\s*parameters\s*\[\s*(\d+)\s*\]\s*\)?\s*$")


class Expr:
    """Immutable expression node.

    ``t`` is a tuple: ('c', value) constant, ('p', i) per-particle parameter
    column i, ('v', name) kernel variable (oodppo, dp, turn), ('r', text) raw
    C expression spliced verbatim, ('+'|'-'|'*'|'/', a, b), ('neg', a),
    ('f', fname, a) unary libm call.
    """

    __slots__ = ("t", "_h")

    def __init__(self, t):
        self.t = t
        self._h = hash(t)

    # construction -------------------------------------------------------
    @staticmethod
    def of(v) -> "Expr":
        if isinstance(v, Expr):
            return v
        if isinstance(v, (bool, np.bool_)):
            return Expr(("c", float(v)))
        if isinstance(v, (int, float, np.integer, np.floating)):
            return Expr(("c", float(v)))
        if isinstance(v, str):
            m = _PARAM_RE.match(v)
            if m:
                return Expr(("p", int(m.group(1))))
            try:
                return Expr(("c", float(v)))
            except ValueError:
                return Expr(("r", v.strip()))
        raise TypeError(f"cannot use {v!r} as a lattice value")

    def __add__(self, o):
        return Expr(("+", self, Expr.of(o)))

    def __radd__(self, o):
        return Expr(("+", Expr.of(o), self))

    def __sub__(self, o):
        return Expr(("-", self, Expr.of(o)))

    def __rsub__(self, o):
        return Expr(("-", Expr.of(o), self))

    def __mul__(self, o):
        return Expr(("*", self, Expr.of(o)))

    def __rmul__(self, o):
        return Expr(("*", Expr.of(o), self))

    def __truediv__(self, o):
        return Expr(("/", self, Expr.of(o)))

    def __rtruediv__(self, o):
        return Expr(("/", Expr.of(o), self))

    def __neg__(self):
        return Expr(("neg", self))

    # identity -----------------------------------------------------------
    def __eq__(self, o):
        return isinstance(o, Expr) and self.t == o.t

    def __hash__(self):
        return self._h

    def __repr__(self):
        return f"Expr{self.t!r}"

    # queries ------------------------------------------------------------
    @property
    def kind(self):
        return self.t[0]

    def is_const(self) -> bool:
        return self.t[0] == "c"

    @property
    def value(self) -> float:
        assert self.t[0] == "c"
        return self.t[1]

    def leaves(self):
        k = self.t[0]
        if k in ("c", "p", "v", "r"):
            yield self
            return
        for child in self.t[1:]:
            if isinstance(child, Expr):
                yield from child.leaves()

    def params(self):
        return {e.t[1] for e in self.leaves() if e.t[0] == "p"}

    def has_raw(self) -> bool:
        return any(e.t[0] == "r" for e in self.leaves())

    def uses_var(self, name) -> bool:
        return any(e.t == ("v", name) for e in self.leaves())


def map_leaves(e: Expr, fn_) -> Expr:
    """Copy of ``e`` with every leaf replaced by ``fn_(leaf)``."""
    t = e.t
    if t[0] in ("c", "p", "v", "r"):
        return fn_(e)
    return Expr(tuple(map_leaves(c, fn_) if isinstance(c, Expr) else c for c in t))


def diff(e: Expr, i: int) -> Expr:
    """Symbolic derivative of ``e`` with respect to parameter column ``i``
    (raw C text and kernel variables are constants)."""
    e = Expr.of(e)
    t = e.t
    k = t[0]
    if k in ("c", "v", "r"):
        return ZERO
    if k == "p":
        return const(1.0) if t[1] == i else ZERO
    if k == "neg":
        return -diff(t[1], i)
    if k == "+":
        return diff(t[1], i) + diff(t[2], i)
    if k == "-":
        return diff(t[1], i) - diff(t[2], i)
    if k == "*":
        return diff(t[1], i) * t[2] + t[1] * diff(t[2], i)
    if k == "/":
        return diff(t[1], i) / t[2] - t[1] * diff(t[2], i) / (t[2] * t[2])
    if k in (">", "<"):
        return ZERO
    if k == "?":
        return select(t[1], diff(t[2], i), diff(t[3], i))
    if k == "f":
        a, da = t[2], diff(t[2], i)
        name = t[1]
        if name == "sqrt":
            return da / (fn("sqrt", a) * 2.0)
        if name == "fabs":
            return select(gt(a, ZERO), da, -da)
        if name == "sin":
            return fn("cos", a) * da
        if name == "cos":
            return -(fn("sin", a) * da)
        if name == "tan":
            return da / (fn("cos", a) * fn("cos", a))
        if name == "sinh":
            return fn("cosh", a) * da
        if name == "cosh":
            return fn("sinh", a) * da
    raise ValueError(f"cannot differentiate {e!r}")


def evaluate(e: Expr, params, env=None, fp32=False) -> float:
    """Value of ``e`` on the host with parameter column i bound to
    ``params[i]`` and kernel variables to ``env``, in FP64 or, with
    ``fp32``, with every leaf and every operation rounded to FP32 (the
    arithmetic of an FP32 kernel evaluating the same expression)."""
    t = Expr.of(e).t
    k = t[0]
    r32 = (lambda v: float(np.float32(v))) if fp32 else (lambda v: v)
    if k == "c":
        return r32(t[1])
    if k == "p":
        return r32(float(params[t[1]]))
    if k == "v":
        return r32(float(env[t[1]]))
    if k == "r":
        raise ValueError("raw C text cannot be evaluated on the host")
    if k == "neg":
        return -evaluate(t[1], params, env, fp32)
    if k == "f":
        return libm(t[1], evaluate(t[2], params, env, fp32), fp32)
    if k == "?":
        return evaluate(t[2] if evaluate(t[1], params, env, fp32) != 0.0 else t[3], params, env, fp32)
    a = evaluate(t[1], params, env, fp32)
    b = evaluate(t[2], params, env, fp32)
    if k == ">":
        return float(a > b)
    if k == "<":
        return float(a < b)
    return _binop(k, a, b, fp32)


def shift_params(e: Expr, offset: int, start: int = 0) -> Expr:
    """Copy of ``e`` with parameter columns >= start moved by ``offset``."""
    return map_leaves(Expr.of(e), lambda l: Expr(("p", l.t[1] + offset)) if l.t[0] == "p" and l.t[1] >= start
                      else l)


def remap_params(e: Expr, mapping) -> Expr:
    """Copy of ``e`` with parameter column i replaced by mapping[i] (columns
    not in the mapping stay)."""
    return map_leaves(Expr.of(e), lambda l: Expr(("p", mapping.get(l.t[1], l.t[1]))) if l.t[0] == "p" else l)


def select(cond, a, b) -> Expr:
    """``cond ? a : b`` (cond is a comparison expression)."""
    return Expr(("?", Expr.of(cond), Expr.of(a), Expr.of(b)))


def gt(a, b) -> Expr:
    return Expr((">", Expr.of(a), Expr.of(b)))


def lt(a, b) -> Expr:
    return Expr(("<", Expr.of(a), Expr.of(b)))


def const(v) -> Expr:
    return Expr(("c", float(v)))


def param(i) -> Expr:
    """Per-particle parameter column ``i`` (a row of the kernel's pool)."""
    return Expr(("p", int(i)))


def var(name) -> Expr:
    return Expr(("v", name))


def fn(name, a) -> Expr:
    return Expr(("f", name, Expr.of(a)))


OODPPO = var("oodppo")
DP = var("dp")
ZERO = const(0.0)


_MP = {"sqrt": mpmath.sqrt, "sin": mpmath.sin, "cos": mpmath.cos, "tan": mpmath.tan,
       "sinh": mpmath.sinh, "cosh": mpmath.cosh}
_LIBM = {"sqrt": math.sqrt, "sin": math.sin, "cos": math.cos, "tan": math.tan,
         "sinh": math.sinh, "cosh": math.cosh}


def libm(name: str, v: float, fp32: bool = False) -> float:
    """``name(v)`` correctly rounded to the kernel precision, as a C
    compiler's constant folder computes it (gcc folds through MPFR): the
    host's libm may be an ulp off, which over many turns of a thick
    element would bias the orbit against a reference that folds exactly."""
    if name == "fabs":
        return abs(v)
    if math.isnan(v) or math.isinf(v) or (name == "sqrt" and v < 0):
        r = _LIBM[name](v) if not (name == "sqrt" and v < 0) else math.nan
        return float(np.float32(r)) if fp32 else r
    with mpmath.workprec(256):
        r = _MP[name](mpmath.mpf(v))
    if fp32:
        with mpmath.workprec(24):
            r = +r
    return float(r)


def fold(e: Expr, fp32: bool = False, env=None) -> Expr:
    """Fold constant sub-trees in the kernel precision.

    ``env`` may bind kernel variables (``oodppo``/``dp``) to constants when
    they are known at code-generation time (a literal ``dp``).
    """
    e = Expr.of(e)
    t = e.t
    k = t[0]
    if k == "c":
        return const(float(np.float32(t[1]))) if fp32 else e
    if k in ("p", "r"):
        return e
    if k == "v":
        if env and t[1] in env:
            return fold(const(env[t[1]]), fp32)
        return e
    if k == "neg":
        a = fold(t[1], fp32, env)
        return const(-a.value) if a.is_const() else Expr(("neg", a))
    if k == "?":
        c = fold(t[1], fp32, env)
        if c.is_const():
            return fold(t[2] if c.value != 0.0 else t[3], fp32, env)
        return Expr(("?", c, fold(t[2], fp32, env), fold(t[3], fp32, env)))
    if k in (">", "<"):
        a = fold(t[1], fp32, env)
        b = fold(t[2], fp32, env)
        if a.is_const() and b.is_const():
            return const(float(a.value > b.value if k == ">" else a.value < b.value))
        return Expr((k, a, b))
    if k == "f":
        a = fold(t[2], fp32, env)
        if a.is_const():
            return const(libm(t[1], float(np.float32(a.value)) if fp32 else a.value, fp32))
        return Expr(("f", t[1], a))
    a = fold(t[1], fp32, env)
    b = fold(t[2], fp32, env)
    if a.is_const() and b.is_const():
        return const(_binop(k, a.value, b.value, fp32))
    # identities with a constant operand; a zero factor is taken as an exact
    # zero term (as it is for every finite value of the other operand)
    if k == "*":
        if is_zero(a) or is_zero(b):
            return const(0.0)
        if a.is_const() and a.value == 1.0:
            return b
        if b.is_const() and b.value == 1.0:
            return a
    elif k == "/":
        if is_zero(a) and not b.is_const():
            return const(0.0)
        if b.is_const() and b.value == 1.0:
            return a
    elif k == "+":
        if is_zero(a):
            return b
        if is_zero(b):
            return a
    elif k == "-":
        if is_zero(b):
            return a
        if is_zero(a):
            return Expr(("neg", b))
    return Expr((k, a, b))


def _binop(k, a, b, fp32):
    with np.errstate(all="ignore"):
        if fp32:
            a, b = np.float32(a), np.float32(b)
        else:
            a, b = np.float64(a), np.float64(b)
        if k == "+":
            r = a + b
        elif k == "-":
            r = a - b
        elif k == "*":
            r = a * b
        elif k == "/":
            r = a / b if b != 0 else (np.float64(np.nan) if a == 0 or a != a else np.copysign(np.inf, a) * np.copysign(1.0, b))
        else:
            raise ValueError(k)
    return float(r)


def is_zero(e: Expr) -> bool:
    return e.is_const() and e.value == 0.0


# --------------------------------------------------------------------------
# Ops
# --------------------------------------------------------------------------


@dataclass(frozen=True)
class Op:
    def exprs(self):
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, Expr):
                yield v
            elif isinstance(v, tuple):
                for u in v:
                    if isinstance(u, Expr):
                        yield u

    def map_exprs(self, fn_):
        kw = {}
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, Expr):
                v = fn_(v)
            elif isinstance(v, tuple):
                v = tuple(fn_(u) if isinstance(u, Expr) else u for u in v)
            kw[f.name] = v
        return type(self)(**kw)


@dataclass(frozen=True)
class Drift(Op):
    """Paraxial drift x += px*L, y += py*L (ufo/methods.py:40-41)."""
    L: Expr


@dataclass(frozen=True)
class ExactDrift(Op):
    """Exact drift, l = L/sqrt((1+dp)^2-px^2-py^2) (ufo/methods.py:29-37)."""
    L: Expr


@dataclass(frozen=True)
class Kick(Op):
    """Thin multipole: w = sum_n (kn[n] + i ks[n]) z^n / n!, z = x + i y;
    px -= Re w, py += Im w, every coefficient times oodppo when ``chroma``
    (ufo/methods.py:45-76)."""
    kn: Tuple[Expr, ...]
    ks: Tuple[Expr, ...]
    chroma: bool


@dataclass(frozen=True)
class Weak(Op):
    """Teapot weak focusing px -= x*w (ufo/methods.py:84)."""
    w: Expr


@dataclass(frozen=True)
class Shift(Op):
    """Misalignment x -= dx, y -= dy (ufo/methods.py:22-26)."""
    dx: Expr
    dy: Expr


@dataclass(frozen=True)
class Edge(Op):
    """Bend pole-face kick (ufo/methods.py:199-211): h = h0 (times oodppo when
    ``chroma``), psi = e - 2 h fringe / cos e (1 + sin^2 e),
    px += x h tan e, py -= y h tan psi."""
    h0: Expr
    e: Expr
    fringe: Expr
    chroma: bool


@dataclass(frozen=True)
class ThickQuad(Op):
    """Thick quadrupole matrices (ufo/methods.py:106-132)."""
    L: Expr
    k1: Expr
    chroma: bool


@dataclass(frozen=True)
class ThickBend(Op):
    """Thick combined-function sector bend with dispersion
    (ufo/methods.py:142-194)."""
    L: Expr
    angle: Expr
    k1: Expr
    chroma: bool


@dataclass(frozen=True)
class Wire(Op):
    """Kick of a wire parallel to the beam at (wx, wy), magnitude k/r
    (ufo/methods.py:230-241)."""
    wx: Expr
    wy: Expr
    k: Expr


@dataclass(frozen=True)
class Aperture(Op):
    """``if (<window> == 0) x = NAN;`` with the window spliced as text, so the
    C precedence of ``== 0`` is the reference's (ufo/elements.py:517-518)."""
    window: str


@dataclass(frozen=True)
class Code(Op):
    """Raw C statements from a user-defined element ``method()``."""
    text: str


@dataclass(frozen=True)
class Observe(Op):
    """Observation point: the state is stored (Track/Optics dumps,
    ufo/track.py:62-70)."""
    pass


# --------------------------------------------------------------------------
# Lowering helpers (the IR form of the reference's pass-method templates)
# --------------------------------------------------------------------------


def E(v) -> Expr:
    return Expr.of(v)


def align(dx, dy):
    """Entry misalignment; the exit one is ``align(-dx, -dy)``
    (ufo/methods.py:22-26, ufo/elements.py:217-218)."""
    return [Shift(E(dx), E(dy))]


def drift(flags, length):
    """ufo/methods.py:28-43."""
    if flags & EXACT:
        return [ExactDrift(E(length))]
    return [Drift(E(length))]


def kick(flags, knl, ksl):
    """ufo/methods.py:45-76: LINEAR keeps orders 0 and 1 only; ACHROMATIC
    drops the oodppo factor.  An empty list contributes nothing."""
    kn = [E(k) for k in knl]
    ks = [E(k) for k in ksl]
    if flags & LINEAR:
        kn = kn[:2]
        ks = ks[:2]
    if not kn and not ks:
        return []
    return [Kick(tuple(kn), tuple(ks), not (flags & ACHROMATIC))]


def teapot(flags, length, slices, knl, ksl, angle=0.0):
    """TEAPOT slicing (ufo/methods.py:79-97): end drifts L/(2(1+n)) (L/2 for
    one slice), n-1 equal inner drifts, n kicks of strength k*L/n, each
    followed by the weak-focusing term angle^2 (L/n) / L / L."""
    slices = int(slices)
    L = E(length)
    a = E(angle)
    if slices == 1:
        outer = L / 2.0
        inner = None
    else:
        outer = L / float(2 * (1 + slices))
        inner = (L - outer * 2.0) / float(slices - 1)
    w = (a * a * L / float(slices)) / L / L
    kn = [E(k) * L / float(slices) for k in knl]
    ks = [E(k) * L / float(slices) for k in ksl]
    ops = list(drift(flags, outer))
    for i in range(slices):
        ops += kick(flags, kn, ks)
        ops.append(Weak(w))
        ops += drift(flags, inner if i < slices - 1 else outer)
    return ops


def pad_multipoles(m, min_order):
    """Pad a multipole list to at least ``min_order`` entries with zeros.
    (The reference appends max(len, min) zeros, ufo/methods.py:16-20; the
    extra trailing zeros change no value and are not kept here.)"""
    m = list(m)
    if len(m) < min_order:
        m += [0.0] * (min_order - len(m))
    return m


def quadrupole(flags, slices, length, k1, dkn, dks):
    """ufo/methods.py:99-134; thick quads ignore dkn/dks."""
    from .flags import KICK
    if flags & KICK:
        knl = pad_multipoles(dkn, 2)
        knl[1] = E(knl[1]) + E(k1)
        return teapot(flags, length, slices, knl, dks)
    return [ThickQuad(E(length), E(k1), not (flags & ACHROMATIC))]


def sbend(flags, slices, length, angle, k1, dkn, dks):
    """ufo/methods.py:136-196."""
    from .flags import KICK
    if flags & KICK:
        knl = pad_multipoles(dkn, 2)
        knl[1] = E(knl[1]) + E(k1)
        return teapot(flags, length, slices, knl, dks, angle)
    return [ThickBend(E(length), E(angle), E(k1), not (flags & ACHROMATIC))]


def edge(flags, length, angle, e, fringe):
    """ufo/methods.py:199-211: curvature = angle/length (times oodppo)."""
    return [Edge(E(angle) / E(length), E(e), E(fringe), not (flags & ACHROMATIC))]


def sextupole(flags, slices, length, k2, k2s, dkn, dks):
    """ufo/methods.py:213-219.  The reference passes the *error* skew list to
    the teapot, so ``k2s`` has no effect; that behaviour is kept."""
    knl = pad_multipoles(dkn, 3)
    knl[2] = E(knl[2]) + E(k2)
    return teapot(flags, length, slices, knl, dks)


def octupole(flags, slices, length, k3, k3s, dkn, dks):
    """ufo/methods.py:221-227 (``k3s`` has no effect, as in the reference)."""
    knl = pad_multipoles(dkn, 4)
    knl[3] = E(knl[3]) + E(k3)
    return teapot(flags, length, slices, knl, dks)


def wire(x, y, k):
    """ufo/methods.py:230-241."""
    return [Wire(E(x), E(y), E(k))]


# --------------------------------------------------------------------------
# Canonical FLOP count (algorithmic work per particle and turn)
# --------------------------------------------------------------------------


def _nz(e: Expr) -> bool:
    return not is_zero(e)


def kick_flops(kn, ks) -> int:
    """Counting rule for a thin kick: with top order N >= 1 (highest non-zero
    coefficient) it is (6 if c_N is complex else 2) + 6 (N-1) + A + 2, where A
    is the number of non-zero real/imaginary parts among c_0..c_{N-1}.  With
    N = 0 it is the number of non-zero parts of c_0."""
    n = max(len(kn), len(ks))
    re = [kn[i] if i < len(kn) else ZERO for i in range(n)]
    im = [ks[i] if i < len(ks) else ZERO for i in range(n)]
    top = -1
    for i in range(n):
        if _nz(re[i]) or _nz(im[i]):
            top = i
    if top < 0:
        return 0
    if top == 0:
        return int(_nz(re[0])) + int(_nz(im[0]))
    cplx = _nz(re[top]) and _nz(im[top])
    a = sum(int(_nz(re[i])) + int(_nz(im[i])) for i in range(top))
    return (6 if cplx else 2) + 6 * (top - 1) + a + 2


def op_flops(op: Op) -> int:
    """Canonical FP64 flops of one op (+,-,* = 1, fma = 2; hoisted invariants
    and zero terms are not counted)."""
    if isinstance(op, Drift):
        return 4 if _nz(op.L) else 0
    if isinstance(op, ExactDrift):
        return 10
    if isinstance(op, Kick):
        return kick_flops(op.kn, op.ks)
    if isinstance(op, Weak):
        return 2 if _nz(op.w) else 0
    if isinstance(op, Shift):
        return int(_nz(op.dx)) + int(_nz(op.dy))
    if isinstance(op, Edge):
        return 4 if (_nz(op.e) or _nz(op.fringe)) and _nz(op.h0) else 0
    if isinstance(op, ThickQuad):
        return 12
    if isinstance(op, ThickBend):
        return 14
    return 0


