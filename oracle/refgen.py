"""ORACLE — test infrastructure only.

Nothing in the product imports this file.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only as the checker.

What it is: a CPU restatement of the reference's lattice-to-kernel generator.
The reference turns a lattice into OpenCL-C text (ufo/methods.py templates,
ufo/elements.py ``method()``, ufo/lattice.py ``Line.method``) and wraps that
text in the kernel bodies of ufo/track.py, ufo/sa.py, ufo/optics.py and
ufo/co.py.  This module emits the same statements, with the same expression
structure and the same literal text for every constant, into a C11
translation unit that ``runner.py`` compiles with gcc and calls through
ctypes.  Built with ``-ffp-contract=off`` it is the reference's arithmetic
under strict IEEE rules; built with ``-ffast-math -fno-finite-math-only
-ffp-contract=fast`` it stands in for the reference's default OpenCL options
``-cl-fast-relaxed-math -cl-mad-enable`` (ufo/__init__.py:22) with NaN
checks kept.

Intentional differences from the reference text (values are unchanged):
  * ``copy_multipoles_and_pad`` pads to ``max(len, min_order)`` entries
    instead of appending that many zeros (ufo/methods.py:16-20).  The extra
    top-order zeros of the reference only add exact zero terms to the Horner
    sum for finite coordinates, and they make the text grow as 2**order.
  * The persistent work pool (``atomic_inc(pool_idx)``) becomes an OpenMP
    ``schedule(dynamic)`` loop; the reference's results do not depend on the
    order in which particles are claimed.

Element objects are read through their attributes only (label, length, k1,
...); the product's own lowering (``ufo_amd.ir``/``codegen``) is never used
here.
"""
from __future__ import annotations

FLAG_LINEAR = 1 << 0
FLAG_FIVED = 1 << 1
FLAG_EXACT = 1 << 2
FLAG_KICK = 1 << 3
FLAG_DOUBLE = 1 << 5
FLAG_ACHROMATIC = 1 << 6


# ---------------------------------------------------------------------------
# Pass-method templates (ufo/methods.py)
# ---------------------------------------------------------------------------

def t_pad(m, min_order):
    """ufo/methods.py:16-20 (see module note on padding)."""
    m = list(m)
    m += ["0."] * max(0, min_order - len(m))
    return m


def t_align(dx, dy):
    """ufo/methods.py:22-26."""
    return f"x -= {dx};\ny -= {dy};\n"


def t_drift(flags, length):
    """ufo/methods.py:28-43."""
    if flags & FLAG_EXACT:
        return ("{\n"
                "    ufloat opdp, l;\n"
                "    opdp = 1. + dp;\n"
                f"    l = {length} / sqrt(opdp*opdp - px*px - py*py);\n"
                "    x += px * l;\n"
                "    y += py * l;\n"
                "}\n")
    return f"x += px * {length};\ny += py * {length};\n"


def _chroma(flags):
    return "" if flags & FLAG_ACHROMATIC else "* oodppo"


def t_kick(flags, knl, ksl):
    """Nested-Horner thin kick, ufo/methods.py:45-76: one update for the
    normal list, one for the skew list, each level dividing by its order."""
    text = ""
    ch = _chroma(flags)
    if knl:
        top = min(len(knl), 2) if flags & FLAG_LINEAR else len(knl)
        re_, im_ = f"{knl[top - 1]} {ch}", "0."
        for order in range(top - 1, 0, -1):
            nxt = f"({re_} * x - ({im_} * y)) / {order}"
            im_ = f"({re_} * y + ({im_} * x)) / {order}"
            re_ = f"({knl[order - 1]} {ch} + {nxt})"
        text = f"px -= {re_};\npy += {im_};\n\n"
    if ksl:
        top = min(len(ksl), 2) if flags & FLAG_LINEAR else len(ksl)
        im_, re_ = f"{ksl[top - 1]} {ch}", "0."
        for order in range(top - 1, 0, -1):
            nxt = f"({re_} * y + ({im_} * x)) / {order}"
            re_ = f"({re_} * x - ({im_} * y)) / {order}"
            im_ = f"({ksl[order - 1]} {ch} + {nxt})"
        text += f"px -= {re_};\npy += {im_};\n\n"
    return text


def t_teapot(flags, length, slices, knl, ksl, angle=0):
    """TEAPOT slicing, ufo/methods.py:79-97."""
    outer = f"({length} / 2.)" if slices == 1 else f"({length} / (2 * (1 + {slices})))"
    inner = f"({length} - 2. * {outer}) / {slices - 1.}"
    weak = f"px -= x * (({angle}) * ({angle}) * ({length})/{slices} / ({length}) / ({length}));\n"
    kn = [f"({k}*{length}/{slices})" for k in knl]
    ks = [f"({k}*{length}/{slices})" for k in ksl]
    text = t_drift(flags, outer)
    for _ in range(slices - 1):
        text += t_kick(flags, kn, ks) + weak
        text += t_drift(flags, inner)
    text += t_kick(flags, kn, ks) + weak
    text += t_drift(flags, outer)
    return text


def t_quadrupole(flags, slices, length, k1, dkn, dks):
    """ufo/methods.py:99-134."""
    if flags & FLAG_KICK:
        knl = t_pad(dkn, 2)
        knl[1] = f"({knl[1]} + {k1})"
        return t_teapot(flags, length, slices, knl, dks)
    return ("{\n"
            "    ufloat x0  = x;\n    ufloat px0 = px;\n    ufloat y0  = y;\n    ufloat py0 = py;\n"
            f"    ufloat k         = {k1};\n"
            f"    ufloat k2        = sqrt(fabs(k {_chroma(flags)}));\n"
            f"    ufloat k2l       = {length} * k2;\n"
            "    ufloat C         = cos( k2l);\n"
            "    ufloat CH        = cosh(k2l);\n"
            "    ufloat S         = sin( k2l);\n"
            "    ufloat SH        = sinh(k2l);\n"
            "    if (k > 0.) {\n"
            "        x  =  C  * x0      + S  * px0 / k2;\n"
            "        px = -S  * x0 * k2 + C  * px0;\n"
            "        y  =  CH * y0      + SH * py0 / k2;\n"
            "        py =  SH * y0 * k2 + CH * py0;\n"
            "    } else {\n"
            "        x  =  CH * x0      + SH * px0 / k2;\n"
            "        px =  SH * x0 * k2 + CH * px0;\n"
            "        y  =  C  * y0      + S  * py0 / k2;\n"
            "        py = -S  * y0 * k2 + C  * py0;\n"
            "   }\n"
            "}\n")


def t_sbend(flags, slices, length, angle, k1, dkn, dks):
    """ufo/methods.py:136-196."""
    if flags & FLAG_KICK:
        knl = t_pad(dkn, 2)
        knl[1] = f"({knl[1]} + {k1})"
        return t_teapot(flags, length, slices, knl, dks, angle)
    ch = _chroma(flags)
    inv = "" if flags & FLAG_ACHROMATIC else "/ oodppo"
    return ("{\n"
            "    ufloat x0  = x;\n    ufloat px0 = px;\n    ufloat y0  = y;\n    ufloat py0 = py;\n"
            f"    ufloat curvature = {angle} / ({length}) {ch};\n"
            f"    ufloat k         = {k1} {ch};\n"
            "    ufloat k2        = sqrt(fabs(k));\n"
            f"    ufloat k2l       = {length} * k2;\n"
            "    ufloat C         = cos( k2l);\n"
            "    ufloat CH        = cosh(k2l);\n"
            "    ufloat S         = sin( k2l);\n"
            "    ufloat SH        = sinh(k2l);\n"
            "    if (k > 0.) {\n"
            "        y  =  CH * y0      + SH * py0 / k2;\n"
            "        py =  SH * y0 * k2 + CH * py0;\n"
            "    }\n"
            "    if (k < 0.) {\n"
            "        y  =  C  * y0      + S  * py0 / k2;\n"
            "        py = -S  * y0 * k2 + C  * py0;\n"
            "    }\n"
            "    if (k == 0.)\n"
            f"        y += {length} * py0;\n"
            f"    k += curvature * curvature {inv};\n"
            "    k2  = sqrt(fabs(k));\n"
            f"    k2l = {length} * k2;\n"
            "    C   = cos( k2l);\n"
            "    CH  = cosh(k2l);\n"
            "    S   = sin( k2l);\n"
            "    SH  = sinh(k2l);\n"
            "    if (k > 0.) {\n"
            "        x   =  C  * x0      + S  * px0 / k2;\n"
            "        x  += dp * curvature * (1. - C) / fabs(k);\n"
            "        px  = -S  * x0 * k2 + C  * px0;\n"
            "        px += dp * curvature * S / k2;\n"
            "    }\n"
            "    if (k < 0.) {\n"
            "        x  =  CH * x0      + SH * px0 / k2;\n"
            "        x +=  dp * curvature * (CH - 1.) / fabs(k);\n"
            "        px =  SH * x0 * k2 + CH * px0;\n"
            "        px += dp * curvature * SH / k2;\n"
            "    }\n"
            "    if (k == 0.)\n"
            f"        x += {length} * px0;\n"
            "}\n")


def t_edge(flags, length, angle, e, fringe):
    """ufo/methods.py:199-211."""
    return ("{\n"
            "    ufloat curvature, psi;\n"
            f"    curvature = {angle} / ({length}) {_chroma(flags)};\n"
            f"    psi = {e} - 2 * curvature * {fringe} / cos({e}) * (1. + sin({e}) * sin({e}));\n"
            f"    px += x * curvature * tan({e});\n"
            "    py -= y * curvature * tan(psi);\n"
            "}\n")


def t_sextupole(flags, slices, length, k2, k2s, dkn, dks):
    """ufo/methods.py:213-219; the teapot gets the error skew list ``dks``."""
    knl = t_pad(dkn, 3)
    knl[2] = f"({knl[2]} + {k2})"
    return t_teapot(flags, length, slices, knl, dks)


def t_octupole(flags, slices, length, k3, k3s, dkn, dks):
    """ufo/methods.py:221-227."""
    knl = t_pad(dkn, 4)
    knl[3] = f"({knl[3]} + {k3})"
    return t_teapot(flags, length, slices, knl, dks)


def t_wire(x, y, k):
    """ufo/methods.py:230-241."""
    return ("{\n"
            "    ufloat alpha, B;\n"
            f"     alpha = atan2({y} - y, {x} - x);\n"
            f"     B = ({k}) / hypot({x} - x, {y} - y);\n"
            "     px += B * cos(alpha);\n"
            "     py += B * sin(alpha);\n"
            "}\n")


# ---------------------------------------------------------------------------
# Element lowering (ufo/elements.py method()) — reads attributes only
# ---------------------------------------------------------------------------

def _dk(kwargs, dkn, dks):
    """ufo/elements.py:14-23."""
    for prm, value in kwargs.items():
        if prm[:2] == "dk" and prm[2:].isdigit():
            i = int(prm[2:])
            dkn += [0.] * (i + 1 - len(dkn))
            dkn[i] = value
        if prm[:2] == "dk" and prm[-1] == "s" and prm[2:-1].isdigit():
            i = int(prm[2:-1])
            dks += [0.] * (i + 1 - len(dks))
            dks[i] = value


def _thick_pieces(fracture, make):
    code = []
    f0 = 0.
    for f in list(fracture) + [1.]:
        code.append(make(f - f0))
        f0 = f
    return code


KINDS = ("Marker", "Drift", "Multipole", "Quadrupole", "Sbend", "Rbend", "Sextupole", "Octupole", "Wire",
         "Aperture")


def element_text(el, flags, fracture, kw):
    """Code pieces of one element, ufo/elements.py:56-518.

    A user subclass of an element class whose ``method()`` calls the base
    method with overrides (the HOWTO's time-dependent kicker,
    HOWTO.md:86-98) is restated by the base class's template with those
    overrides, which the subclass lists as ``oracle_overrides`` (name ->
    the C text its method passes)."""
    kind = next((c.__name__ for c in type(el).__mro__ if c.__name__ in KINDS), type(el).__name__)
    extra = getattr(el, "oracle_overrides", None)
    if extra:
        kw = {**kw, **extra}
    g = lambda name: kw.get(name) or getattr(el, name)   # noqa: E731  (v or self.v)
    if kind == "Marker":
        return [""]
    if kind == "Drift":
        length = g("length")
        return _thick_pieces(fracture, lambda fr: t_drift(flags, f"{length} * {fr}"))
    if kind == "Multipole":
        knl = list(el.knl) if kw.get("knl") is None else list(kw["knl"])
        ksl = list(el.ksl) if kw.get("ksl") is None else list(kw["ksl"])
        dx, dy = g("dx"), g("dy")
        for prm, value in kw.items():
            if prm[0] == "k" and prm[1:].isdigit():
                knl[int(prm[1:])] = value
            if prm[0] == "k" and prm[-1] == "s" and prm[1:-1].isdigit():
                ksl[int(prm[1:-1])] = value
        return [t_align(dx, dy) + t_kick(flags, knl, ksl) + t_align(f"-({dx})", f"-({dy})")]
    if kind in ("Quadrupole", "Sbend", "Rbend", "Sextupole", "Octupole"):
        slices = int(g("slices"))
        length = g("length")
        dx, dy = g("dx"), g("dy")
        dkn = list(el.dkn) if kw.get("dkn") is None else list(kw["dkn"])
        dks = list(el.dks) if kw.get("dks") is None else list(kw["dks"])
        _dk(kw, dkn, dks)
        if kind == "Quadrupole":
            k1 = g("k1")
            code = _thick_pieces(fracture, lambda fr: t_quadrupole(flags, slices, f"{length} * {fr}", k1, dkn, dks))
            code[0] = t_align(dx, dy) + code[0]
            code[-1] = code[-1] + t_align(f"-({dx})", f"-({dy})")
            return code
        if kind in ("Sbend", "Rbend"):
            angle, k1 = g("angle"), g("k1")
            e1, e2, hgap, fint = g("e1"), g("e2"), g("hgap"), g("fint")
            code = _thick_pieces(fracture, lambda fr: t_sbend(flags, slices, f"{length} * {fr}",
                                                              f"{angle} * {fr}", k1, dkn, dks))
            fringe = f"({hgap} * {fint})"
            if kind == "Rbend":
                e1 = f"{e1} + ({angle}*0.5)"
                e2 = f"{e2} + ({angle}*0.5)"
            code[0] = t_align(dx, dy) + t_edge(flags, length, angle, e1, fringe) + code[0]
            code[-1] += t_edge(flags, length, angle, e2, fringe) + t_align(f"-({dx})", f"-({dy})")
            return code
        if kind == "Sextupole":
            k2, k2s = g("k2"), g("k2s")
            code = _thick_pieces(fracture, lambda fr: t_sextupole(flags, slices, f"{length} * {fr}", k2, k2s, dkn, dks))
        else:
            k3, k3s = g("k3"), g("k3s")
            code = _thick_pieces(fracture, lambda fr: t_octupole(flags, slices, f"{length} * {fr}", k3, k3s, dkn, dks))
        code[0] = t_align(dx, dy) + code[0]
        code[-1] = code[-1] + t_align(f"-({dx})", f"-({dy})")
        return code
    if kind == "Wire":
        return [t_wire(g("x"), g("y"), g("k"))]
    if kind == "Aperture":
        return [f"if ({el.window} == 0) x = NAN;\n"]
    raise NotImplementedError(f"oracle has no template for {kind}")


def flatten(line):
    out = []
    for e in line:
        if getattr(e, "flatten", None):
            out += flatten(e)
        else:
            out.append(e)
    return out


def line_text(line, fracture=(), flags=0, parameters=()):
    """ufo/lattice.py:86-119: code pieces split at fracture points, with the
    user's parameters bound to ``parameters[i]``."""
    code = [""]
    for idx, el in enumerate(flatten(line)):
        where = [f - idx for f in fracture if f != "*" and idx <= f < idx + 1]
        if 0 in where:
            where.remove(0)
            code.append("")
        kw = {}
        for i, p in enumerate(parameters):
            if p[0] == el.label or p[0] == idx:
                kw[p[1]] = f"parameters[{i}]"
        pieces = element_text(el, flags, where, kw)
        code[-1] += pieces.pop(0)
        code += pieces
    return code


# ---------------------------------------------------------------------------
# Kernel bodies as C11 functions
# ---------------------------------------------------------------------------

# The reference calls its platform's sin, cos, tan, sinh, cosh (OpenCL
# builtins, whose rounding the standard leaves open).  The restatement
# evaluates them in long double and rounds once to the kernel precision:
# correctly rounded but for ~3e-4 of arguments (those within ~2^-11 ulp of a
# rounding boundary; checked against mpmath), where the host's double sinh and
# cosh are an ulp off for about a quarter of all arguments.  A gcc constant
# fold of the same call (a literal dp) is correctly rounded through MPFR.
_PRELUDE = """#include <tgmath.h>
#include <math.h>
#include <stddef.h>
typedef {ufloat} ufloat;
#undef sin
#undef cos
#undef tan
#undef sinh
#undef cosh
#define sin(v) ((ufloat)sinl((long double)(v)))
#define cos(v) ((ufloat)cosl((long double)(v)))
#define tan(v) ((ufloat)tanl((long double)(v)))
#define sinh(v) ((ufloat)sinhl((long double)(v)))
#define cosh(v) ((ufloat)coshl((long double)(v)))
"""


def _init_coords(parameters, fived, dp):
    """Coordinate initialisation of ufo/track.py:93-103 / ufo/sa.py:101-111."""
    text = ""
    for p in ["x", "px", "y", "py", "z", "dp"]:
        src = f"parameters[{list(parameters).index(p)}]" if p in parameters else "0"
        text += f"        {p}  = {src};\n"
    if fived:
        text += "        z  = 0.;\n"
        text += f"        dp = {dp};\n"
    text += "        oodppo = 1. / (dp + 1.);\n"
    return text


def track_source(line, flags=0, turns=1, parameters=(), where=(), dp=0.):
    """Track kernel (ufo/track.py:56-114) as ``run(pool, P, particles,
    tracks)`` with tracks laid out ``[particles][count*turns][6]``.
    Returns (source, count*turns, dumps per turn)."""
    fived = flags & FLAG_FIVED
    where = list(where) if isinstance(where, (list, tuple)) else [where]
    dump = ("tr[track_idx][0] = x;\ntr[track_idx][1] = px;\n"
            "tr[track_idx][2] = y;\ntr[track_idx][3] = py;\n")
    if not fived:
        dump += "tr[track_idx][4] = z;\ntr[track_idx][5] = dp;\n"
    dump += "track_idx++;\n\n"
    pieces = line_text(line, fracture=where, flags=flags, parameters=parameters)
    body = dump.join(pieces)
    ndump = len(pieces) - 1
    if -1 in where:
        body += dump
        ndump += 1
    count = len(where) * turns
    ufloat = "double" if flags & FLAG_DOUBLE else "float"
    src = _PRELUDE.format(ufloat=ufloat)
    src += ("void run(const ufloat *pool, int P, int particles, ufloat *tracks)\n{\n"
            "    #pragma omp parallel for schedule(dynamic, 1)\n"
            "    for (int idx = 0; idx < particles; idx++) {\n"
            "        const ufloat *parameters = pool + (size_t)idx * P;\n"
            "        ufloat x, y, z, px, py, dp, oodppo;\n"
            "        int track_idx = 0;\n"
            f"        ufloat (*tr)[6] = (ufloat (*)[6])(tracks + (size_t)idx * {count} * 6);\n")
    src += _init_coords(parameters, fived, dp)
    src += (f"        for (int turn = 0; turn < {turns}; turn++) {{\n"
            + body +
            "            if (isnan(x))\n                break;\n"
            "        }\n"
            "        (void)track_idx; (void)z; (void)px; (void)py; (void)oodppo;\n"
            "    }\n}\n")
    return src, count, ndump


def sa_source(line, flags=0, turns=1000, parameters=(), dp=0.):
    """StableAperture kernel (ufo/sa.py:52-126) as ``run(pool, P, particles,
    lost)``: lost[i] = completed turns before the end-of-turn loss test
    ``fabs(x) > 1 || fabs(y) > 1 || isnan(x)`` fired, or ``turns``."""
    fived = flags & FLAG_FIVED
    body = "".join(line_text(line, flags=flags, parameters=parameters))
    ufloat = "double" if flags & FLAG_DOUBLE else "float"
    src = _PRELUDE.format(ufloat=ufloat)
    src += ("void run(const ufloat *pool, int P, int particles, unsigned int *lost)\n{\n"
            "    #pragma omp parallel for schedule(dynamic, 1)\n"
            "    for (int idx = 0; idx < particles; idx++) {\n"
            "        const ufloat *parameters = pool + (size_t)idx * P;\n"
            "        ufloat x, y, z, px, py, dp, oodppo;\n"
            "        int turn;\n")
    src += _init_coords(parameters, fived, dp)
    src += (f"        for (turn = 0; turn < {turns}; turn++) {{\n"
            + body +
            "            if (fabs(x) > 1. || fabs(y) > 1. || isnan(x))\n                break;\n"
            "        }\n"
            "        lost[idx] = turn;\n"
            "        (void)z; (void)px; (void)py; (void)oodppo;\n"
            "    }\n}\n")
    return src


def optics_source(line, where=(), periodic=True, parameters=(), flags=FLAG_LINEAR | FLAG_ACHROMATIC):
    """Optics kernel (ufo/optics.py:53-115): three rays x=y=1, px=py=1, dp=1
    from pool row 0, dumps at the observation points.  Returns (source,
    count, where) with ``where`` after the periodic end-of-line append."""
    where = list(where) if isinstance(where, (list, tuple)) else [where]
    if periodic and -1 not in where and line[-1] not in where:
        where.append(-1)
    dump = ("tr[track_idx][0] = x;\ntr[track_idx][1] = px;\n"
            "tr[track_idx][2] = y;\ntr[track_idx][3] = py;\ntrack_idx++;\n\n")
    pieces = line_text(line, fracture=where, flags=flags, parameters=parameters)
    body = dump.join(pieces)
    count = len(where) - 1
    if -1 in where:
        count += 1
        body += dump
    ufloat = "double" if flags & FLAG_DOUBLE else "float"
    src = _PRELUDE.format(ufloat=ufloat)
    src += ("void run(const ufloat *pool, int P, int particles, ufloat *tracks)\n{\n"
            "    for (int particle_idx = 0; particle_idx < 3; particle_idx++) {\n"
            "        const ufloat *parameters = pool;\n"
            "        ufloat x, y, z, px, py, dp, oodppo;\n"
            "        int track_idx = 0;\n"
            f"        ufloat (*tr)[6] = (ufloat (*)[6])(tracks + (size_t)particle_idx * {count} * 6);\n"
            "        x = y = z = px = py = dp = 0.;\n"
            "        if (particle_idx == 0) {x  = 1.; y  = 1.;}\n"
            "        if (particle_idx == 1) {px = 1.; py = 1.;}\n"
            "        if (particle_idx == 2) dp = 1.;\n")
    if "dp" in parameters:
        src += f"        dp += parameters[{list(parameters).index('dp')}];\n"
    src += ("        oodppo = 1. / (dp + 1.);\n"
            + body +
            "        (void)track_idx; (void)z; (void)oodppo; (void)parameters;\n"
            "    }\n}\n")
    return src, count, where


_AMOEBA = """
#define ALPHA 1.0
#define BETA 0.5
#define GAMMA 2.0
static void simplex_sum(ufloat s[5][4], ufloat *sum)
{
    for (int j = 0; j < 4; j++) {
        sum[j] = 0.0;
        for (int i = 0; i < 5; i++)
            sum[j] += s[i][j];
    }
}
static ufloat amoeba_try(ufloat s[5][4], ufloat *y, ufloat *sum, int ihi, ufloat fac, const ufloat *parameters)
{
    ufloat t[4];
    ufloat fac1 = (1.0 - fac) / 4;
    ufloat fac2 = fac1 - fac;
    for (int j = 0; j < 4; j++)
        t[j] = sum[j] * fac1 - s[ihi][j] * fac2;
    ufloat y_try = method(t, parameters);
    if (y_try < y[ihi]) {
        y[ihi] = y_try;
        for (int j = 0; j < 4; j++) {
            sum[j] += t[j] - s[ihi][j];
            s[ihi][j] = t[j];
        }
    }
    return y_try;
}
static int amoeba(ufloat s[5][4], const ufloat *parameters)
{
    int ilo, ihi, inhi, iter = 0;
    ufloat y_try, y_save, y[5], psum[4];
    for (int i = 0; i < 5; i++)
        y[i] = method(s[i], parameters);
    simplex_sum(s, psum);
    for (;;) {
        ilo = 1;
        ihi = y[1] > y[2] ? (inhi = 2, 1) : (inhi = 1, 2);
        for (int i = 0; i < 5; i++) {
            if (y[i] < y[ilo])
                ilo = i;
            if (y[i] > y[ihi]) {
                inhi = ihi;
                ihi = i;
            } else if (y[i] > y[inhi])
                if (i != ihi)
                    inhi = i;
        }
        if (iter >= ITERATIONS)
            break;
        y_try = amoeba_try(s, y, psum, ihi, -ALPHA, parameters);
        iter++;
        if (y_try <= y[ilo])
            y_try = amoeba_try(s, y, psum, ihi, GAMMA, parameters);
        else if (y_try >= y[inhi]) {
            y_save = y[ihi];
            y_try = amoeba_try(s, y, psum, ihi, BETA, parameters);
            if (y_try >= y_save) {
                for (int i = 0; i < 5; i++) {
                    if (i != ilo) {
                        for (int j = 0; j < 4; j++) {
                            psum[j] = s[i][j] + s[ilo][j];
                            psum[j] *= 0.5;
                            s[i][j] = psum[j];
                        }
                        y[i] = method(psum, parameters);
                    }
                }
                iter += 4;
                simplex_sum(s, psum);
            }
        }
    }
    return ilo;
}
"""


def co_source(line, flags=0, parameters=(), iterations=200):
    """ClosedOrbit kernel (ufo/co.py:146-231): per particle a 4-D
    Nelder-Mead ("amoeba", ufo/co.py:18-124) on the one-turn squared
    residual with dp = 0 and oodppo = 1 (ufo/co.py:175-193)."""
    body = "".join(line_text(line, flags=flags, parameters=parameters))
    ufloat = "double" if flags & FLAG_DOUBLE else "float"
    src = _PRELUDE.format(ufloat=ufloat)
    src += f"#define ITERATIONS {iterations}\n"
    src += ("static ufloat method(const ufloat orbit[4], const ufloat *parameters)\n{\n"
            "    ufloat x = orbit[0], px = orbit[1], y = orbit[2], py = orbit[3];\n"
            "    ufloat z = 0., dp = 0., oodppo = 1.;\n"
            + body +
            "    ufloat residuals;\n"
            "    residuals  = (orbit[0] - x ) * (orbit[0] - x );\n"
            "    residuals += (orbit[1] - px) * (orbit[1] - px);\n"
            "    residuals += (orbit[2] - y ) * (orbit[2] - y );\n"
            "    residuals += (orbit[3] - py) * (orbit[3] - py);\n"
            "    (void)z; (void)dp; (void)oodppo; (void)parameters;\n"
            "    return residuals;\n}\n")
    src += _AMOEBA
    src += ("void run(const ufloat *pool, int P, int particles, ufloat *orbits)\n{\n"
            "    #pragma omp parallel for schedule(dynamic, 1)\n"
            "    for (int idx = 0; idx < particles; idx++) {\n"
            "        const ufloat *parameters = pool + (size_t)idx * P;\n"
            "        ufloat s[5][4];\n")
    for i, p in enumerate(["x", "px", "y", "py"]):
        v = f"parameters[{list(parameters).index(p)}]" if p in parameters else "0."
        for k in range(5):
            src += f"        s[{k}][{i}] = {v};\n"
    src += ("        s[1][0] += 1e-4;\n        s[2][1] += 1e-4;\n"
            "        s[3][2] += 1e-4;\n        s[4][3] += 1e-4;\n"
            "        int ilo = amoeba(s, parameters);\n"
            "        for (int j = 0; j < 4; j++)\n"
            "            orbits[(size_t)idx * 4 + j] = s[ilo][j];\n"
            "    }\n}\n")
    return src
