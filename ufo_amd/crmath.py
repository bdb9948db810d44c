"""Correctly rounded sin, cos, tan, sinh, cosh for the per-particle
invariants of the generated kernels.

A thick element's matrix coefficients depend on 1/(1+dp) when dp is a
per-particle coordinate, so the kernel evaluates their transcendental
functions once per particle.  The device library returns results within an
ulp or so; the reference's arithmetic (a C library, or a compiler folding
through MPFR) is correctly rounded almost always.  A coefficient one ulp off
is a per-particle systematic difference that grows linearly over the turns,
so these invariants are evaluated in double-double arithmetic (Dekker's
TwoSum and FMA-based TwoProduct) and rounded once:

  sin, cos  x = n pi/2 + r, with pi/2 in three parts (each product n*P_i
            exact as a double-double), |r| <= pi/4; Taylor series of sin r
            and cos r to 2^-105, quadrant by n mod 4.
  sinh, cosh  |x| <= 1: Taylor series; otherwise exp(|x|) = 2^k exp(r),
            r = |x| - k ln 2 (ln 2 in three parts), Taylor series of exp r,
            then (E +- 1/E) / 2 with a double-double quotient.
  tan       sin / cos as a double-double quotient.

The series coefficients are 1/k! as double-double constants computed here
with mpmath.  Results are the double nearest the double-double value, i.e.
correctly rounded unless the exact result lies within ~2^-100 relative of a
rounding boundary.  The code uses only +, -, *, / and fma, so the kernels
must be built with -ffp-contract=off (they are), and the host build of the
same text gives the same bits.  The functions are inlined: out-of-line
calls from kernels at the register limit have given wrong results on the
device.  FP32 kernels round the double result to float.
Invariants are evaluated once per particle; inside turn loops the device
library's functions are used.
"""
from __future__ import annotations

import mpmath

FUNCS = ("sin", "cos", "tan", "sinh", "cosh")


def _dd(v) -> str:
    """A value as a double-double literal pair (hi, lo)."""
    with mpmath.workprec(300):
        v = mpmath.mpf(v)
        hi = float(v)
        lo = float(v - mpmath.mpf(hi))
    return f"{{{hi!r}, {lo!r}}}"


def _split3(value):
    """value() (evaluated at 400 bits) as three doubles p1 + p2 + p3, p1 with
    30 significant bits (so that n * p1 is exact for |n| < 2^23), p2 and p3
    the rounded remainders."""
    with mpmath.workprec(400):
        v = mpmath.mpf(value())
        m, e = mpmath.frexp(v)
        p1 = mpmath.ldexp(mpmath.floor(mpmath.ldexp(m, 30)), e - 30)
        r = v - p1
        p2 = mpmath.mpf(float(r))
        p3 = mpmath.mpf(float(r - p2))
        return float(p1), float(p2), float(p3)


def _coeffs(kind: str, n: int, alternate: bool = False):
    """Taylor coefficients 1/k! of the odd (sin, sinh), even (cos, cosh) or
    all (exp) powers, highest power first; ``alternate`` gives the i-th
    term the sign (-1)^i (sin, cos)."""
    with mpmath.workprec(300):
        if kind == "odd":
            ks = [2 * i + 1 for i in range(n)]
        elif kind == "even":
            ks = [2 * i for i in range(n)]
        else:
            ks = list(range(n))
        cs = [(-1 if alternate and i % 2 else 1) / mpmath.factorial(k) for i, k in enumerate(ks)]
        return [_dd(c) for c in reversed(cs)]


def source() -> str:
    """C++/HIP text of the double-double functions (``__device__`` is
    defined empty by a host build)."""
    p1, p2, p3 = _split3(lambda: mpmath.pi / 2)
    l1, l2, l3 = _split3(lambda: mpmath.log(2))
    with mpmath.workprec(100):
        two_over_pi = float(2 / mpmath.pi)
        inv_ln2 = float(1 / mpmath.log(2))
    sin_c = ",\n    ".join(_coeffs("odd", 15, True))    # |r| <= pi/4: r^31/31! < 2^-110
    cos_c = ",\n    ".join(_coeffs("even", 15, True))   # r^30/30! < 2^-105
    sinh_c = ",\n    ".join(_coeffs("odd", 17))         # |x| <= 1: x^35/35! < 2^-130
    cosh_c = ",\n    ".join(_coeffs("even", 17))        # x^34/34! < 2^-120
    exp_c = ",\n    ".join(_coeffs("all", 25))          # |r| <= ln2/2: r^25/25! < 2^-120
    return f"""
// correctly rounded sin/cos/tan/sinh/cosh of per-particle invariants
// (double-double evaluation, see ufo_amd/crmath.py)
struct ufo_dd {{ double hi, lo; }};
__device__ static inline ufo_dd ufo_dd_two_sum(double a, double b) {{
    double s = a + b; double bb = s - a; return {{s, (a - (s - bb)) + (b - bb)}};
}}
__device__ static inline ufo_dd ufo_dd_fast(double a, double b) {{
    double s = a + b; return {{s, b - (s - a)}};
}}
__device__ static inline ufo_dd ufo_dd_add(ufo_dd x, ufo_dd y) {{
    ufo_dd s = ufo_dd_two_sum(x.hi, y.hi); ufo_dd t = ufo_dd_two_sum(x.lo, y.lo);
    s.lo += t.hi; s = ufo_dd_fast(s.hi, s.lo); s.lo += t.lo; return ufo_dd_fast(s.hi, s.lo);
}}
__device__ static inline ufo_dd ufo_dd_mul(ufo_dd x, ufo_dd y) {{
    double p = x.hi * y.hi; double e = fma(x.hi, y.hi, -p);
    e = e + (x.hi * y.lo + x.lo * y.hi); return ufo_dd_fast(p, e);
}}
__device__ static inline ufo_dd ufo_dd_mul_d(ufo_dd x, double d) {{
    double p = x.hi * d; double e = fma(x.hi, d, -p);
    e = e + x.lo * d; return ufo_dd_fast(p, e);
}}
__device__ static inline ufo_dd ufo_dd_neg(ufo_dd x) {{ return {{-x.hi, -x.lo}}; }}
__device__ static inline ufo_dd ufo_dd_div(ufo_dd a, ufo_dd b) {{
    double q1 = a.hi / b.hi; ufo_dd r = ufo_dd_add(a, ufo_dd_neg(ufo_dd_mul_d(b, q1)));
    double q2 = r.hi / b.hi; r = ufo_dd_add(r, ufo_dd_neg(ufo_dd_mul_d(b, q2)));
    double q3 = r.hi / b.hi; ufo_dd q = ufo_dd_fast(q1, q2); return ufo_dd_add(q, {{q3, 0.0}});
}}
__device__ static inline ufo_dd ufo_dd_horner(const ufo_dd *c, int n, ufo_dd z) {{
    ufo_dd s = c[0];
    for (int i = 1; i < n; ++i) s = ufo_dd_add(ufo_dd_mul(s, z), c[i]);
    return s;
}}
__device__ static const ufo_dd ufo_dd_sin_c[15] = {{
    {sin_c}}};
__device__ static const ufo_dd ufo_dd_cos_c[15] = {{
    {cos_c}}};
__device__ static const ufo_dd ufo_dd_sinh_c[17] = {{
    {sinh_c}}};
__device__ static const ufo_dd ufo_dd_cosh_c[17] = {{
    {cosh_c}}};
__device__ static const ufo_dd ufo_dd_exp_c[25] = {{
    {exp_c}}};
// x - n c with c = c1 + c2 + c3, every product exact as a double-double
__device__ static inline ufo_dd ufo_dd_reduce(double x, double n, double c1, double c2, double c3) {{
    double p = n * c1; ufo_dd r = ufo_dd_two_sum(x, -p); r.lo -= fma(n, c1, -p);
    r = ufo_dd_fast(r.hi, r.lo);
    double q = n * c2; r = ufo_dd_add(r, {{-q, -fma(n, c2, -q)}});
    return ufo_dd_add(r, {{-(n * c3), 0.0}});
}}
__device__ static inline void ufo_dd_sincos(double x, ufo_dd *s, ufo_dd *c) {{
    double n = rint(x * {two_over_pi!r});
    ufo_dd r = ufo_dd_reduce(x, n, {p1!r}, {p2!r}, {p3!r});
    ufo_dd r2 = ufo_dd_mul(r, r);
    ufo_dd sr = ufo_dd_mul(ufo_dd_horner(ufo_dd_sin_c, 15, r2), r);
    ufo_dd cr = ufo_dd_horner(ufo_dd_cos_c, 15, r2);
    int q = ((int)fmod(n, 4.0) + 4) & 3;
    if (q == 0) {{ *s = sr; *c = cr; }}
    else if (q == 1) {{ *s = cr; *c = ufo_dd_neg(sr); }}
    else if (q == 2) {{ *s = ufo_dd_neg(sr); *c = ufo_dd_neg(cr); }}
    else {{ *s = ufo_dd_neg(cr); *c = sr; }}
}}
__device__ static inline ufo_dd ufo_dd_exp(double x) {{
    double k = rint(x * {inv_ln2!r});
    ufo_dd r = ufo_dd_reduce(x, k, {l1!r}, {l2!r}, {l3!r});
    ufo_dd e = ufo_dd_horner(ufo_dd_exp_c, 25, r);
    return {{ldexp(e.hi, (int)k), ldexp(e.lo, (int)k)}};
}}
__device__ static inline double ufo_cr_sin(double x) {{
    if (!(fabs(x) < 1e6)) return sin(x);
    ufo_dd s, c; ufo_dd_sincos(x, &s, &c); return s.hi + s.lo;
}}
__device__ static inline double ufo_cr_cos(double x) {{
    if (!(fabs(x) < 1e6)) return cos(x);
    ufo_dd s, c; ufo_dd_sincos(x, &s, &c); return c.hi + c.lo;
}}
__device__ static inline double ufo_cr_tan(double x) {{
    if (!(fabs(x) < 1e6)) return tan(x);
    ufo_dd s, c; ufo_dd_sincos(x, &s, &c); ufo_dd t = ufo_dd_div(s, c); return t.hi + t.lo;
}}
__device__ static inline double ufo_cr_cosh(double x) {{
    double a = fabs(x);
    if (!(a < 700.0)) return cosh(x);
    if (a <= 1.0) {{ ufo_dd z = {{x, 0.0}}; ufo_dd c = ufo_dd_horner(ufo_dd_cosh_c, 17, ufo_dd_mul(z, z)); return c.hi + c.lo; }}
    ufo_dd e = ufo_dd_exp(a); ufo_dd c = ufo_dd_add(e, ufo_dd_div({{1.0, 0.0}}, e));
    return 0.5 * c.hi + 0.5 * c.lo;
}}
__device__ static inline double ufo_cr_sinh(double x) {{
    double a = fabs(x);
    if (!(a < 700.0)) return sinh(x);
    if (a <= 1.0) {{
        ufo_dd z = {{x, 0.0}};
        ufo_dd s = ufo_dd_mul(ufo_dd_horner(ufo_dd_sinh_c, 17, ufo_dd_mul(z, z)), z); return s.hi + s.lo;
    }}
    ufo_dd e = ufo_dd_exp(a); ufo_dd s = ufo_dd_add(e, ufo_dd_neg(ufo_dd_div({{1.0, 0.0}}, e)));
    double r = 0.5 * s.hi + 0.5 * s.lo;
    return x < 0.0 ? -r : r;
}}
"""


def call(name: str, arg: str, fp32: bool) -> str:
    """C text of the correctly rounded call (FP32: the double result rounded
    to float)."""
    if fp32:
        return f"((float)ufo_cr_{name}((double)({arg})))"
    return f"ufo_cr_{name}({arg})"


def attach(src: str) -> str:
    """``src`` with the functions inserted after its prelude when it calls
    them."""
    if "ufo_cr_" not in src or "struct ufo_dd" in src:
        return src
    i = src.find("#endif\n")
    i = i + len("#endif\n") if i >= 0 else 0
    return src[:i] + source() + src[i:]
