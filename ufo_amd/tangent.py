"""Forward-tangent (JVP) emission: the state and K tangent columns
(dx, dpx, dy, dpy) per particle, every op's Jacobian applied to the columns.

Every op's Jacobian depends only on coordinates the op itself does not
change (a kick reads x, y and changes px, py; a drift reads px, py and
changes x, y), so each op updates the state and then the columns with the
same pre-op values.  ``z`` and ``dp`` are constants of the motion (no
element changes them), so the 4x4 block is the whole transverse Jacobian at
fixed dp.

Kernels:
  ``ufo_jvp``   track ``turns`` turns from the pool and write state and
                columns: out[particle][4 + 4K] (x, px, y, py, then column k
                as dx_k, dpx_k, dy_k, dpy_k), initial columns = identity.
  ``ufo_co``    closed orbit by Newton per particle: X <- X - (J - I)^-1
                (F(X) - X) with F the one-turn map and J its Jacobian from the
                tangent columns, until max |F(X) - X| <= tol or ``iterations``.
                dp = 0 and oodppo = 1 as in the reference's objective
                (ufo/co.py:175-193).
"""
from __future__ import annotations

from typing import List

from . import codegen, ir
from .lattice import lower
from .codegen import Env, _fma, _fma_neg, is_zero, emit_op, kick_coeffs

def _loop(body: str) -> str:
    return f"_Pragma(\"unroll\") for (int k = 0; k < NT; ++k) {{ {body} }}"


def emit_tangent(op, env: Env, out: List[str]):
    """Column update of one op (the state update is ``codegen.emit_op``)."""
    f = env.f
    if isinstance(op, ir.Drift):
        L = f(op.L)
        if is_zero(L):
            return
        v = env.val(L)
        out.append(_loop(f"tx[k] = fma(tpx[k], {v}, tx[k]); ty[k] = fma(tpy[k], {v}, ty[k]);"))
    elif isinstance(op, ir.ExactDrift):
        # x += px l, l = L / pz, pz = sqrt((1+dp)^2 - px^2 - py^2):
        # dx = l dpx + px L (px dpx + py dpy) / pz^3
        L = env.val(op.L)
        opdp2 = env.val(ir.E(1.0 + ir.DP) * (1.0 + ir.DP))
        a, b, c = env.new_tmp(), env.new_tmp(), env.new_tmp()
        out.append(f"{{ ufloat {a} = {env.lit(1.0)} / sqrt({_fma('-py', 'py', _fma('-px', 'px', opdp2))}); "
                   f"ufloat {b} = {L} * {a}; ufloat {c} = {b} * {a} * {a}; "
                   + _loop(f"ufloat g = {c} * fma(px, tpx[k], py * tpy[k]); "
                           f"tx[k] = fma(tpx[k], {b}, fma(px, g, tx[k])); "
                           f"ty[k] = fma(tpy[k], {b}, fma(py, g, ty[k]));") + " }")
    elif isinstance(op, ir.Kick):
        emit_kick_tangent(op, env, out)
    elif isinstance(op, ir.Weak):
        w = f(op.w)
        if not is_zero(w):
            out.append(_loop(f"tpx[k] = fma({env.neg(w)}, tx[k], tpx[k]);"))
    elif isinstance(op, ir.Edge):
        h = op.h0 * ir.OODPPO if op.chroma else op.h0
        t1 = f(h * ir.fn("tan", op.e))
        psi = op.e - h * 2.0 * op.fringe / ir.fn("cos", op.e) * (ir.E(1.0) + ir.fn("sin", op.e) * ir.fn("sin", op.e))
        t2 = f(h * ir.fn("tan", psi))
        body = ""
        if not is_zero(t1):
            body += f"tpx[k] = fma(tx[k], {env.val(t1)}, tpx[k]); "
        if not is_zero(t2):
            body += f"tpy[k] = fma(-ty[k], {env.val(t2)}, tpy[k]); "
        if body:
            out.append(_loop(body))
    elif isinstance(op, (ir.ThickQuad, ir.ThickBend)):
        emit_thick_tangent(op, env, out)
    elif isinstance(op, ir.Wire):
        # f = k a / r^2, g = k b / r^2 with a = wx - x, b = wy - y
        wx, wy, kk = env.val(op.wx), env.val(op.wy), env.val(op.k)
        out.append(f"{{ ufloat a = {wx} - x, b = {wy} - y; ufloat r2 = fma(a, a, b * b); "
                   f"ufloat q = {kk} / (r2 * r2); ufloat faa = q * fma(b, b, -a * a); "
                   f"ufloat fab = {env.lit(-2.0)} * q * a * b; "
                   + _loop("tpx[k] = fma(-faa, tx[k], fma(-fab, ty[k], tpx[k])); "
                           "tpy[k] = fma(-fab, tx[k], fma(faa, ty[k], tpy[k]));") + " }")
    elif isinstance(op, (ir.Shift, ir.Aperture, ir.Observe)):
        return
    elif isinstance(op, ir.Code):
        raise NotImplementedError("raw user code has no Jacobian: tangent kernels need IR elements")
    else:
        raise TypeError(f"no tangent for {op!r}")


def emit_kick_tangent(op: ir.Kick, env: Env, out: List[str]):
    """w'(z) = sum n d_n z^(n-1) by Horner; columns: dpx -= Re(w' dz),
    dpy += Im(w' dz)."""
    d = kick_coeffs(op, env)
    top = -1
    for i, (cr, ci) in enumerate(d):
        if not is_zero(cr) or not is_zero(ci):
            top = i
    if top < 1:
        return
    e = [(env.f(d[n + 1][0] * float(n + 1)), env.f(d[n + 1][1] * float(n + 1))) for n in range(top)]
    lines = []
    wr = None if is_zero(e[-1][0]) else env.val(e[-1][0])
    wi = None if is_zero(e[-1][1]) else env.val(e[-1][1])
    for lvl in range(top - 2, -1, -1):
        er = None if is_zero(e[lvl][0]) else env.val(e[lvl][0])
        ei = None if is_zero(e[lvl][1]) else env.val(e[lvl][1])
        r = codegen._chain([(wr, "x", False), (wi, "y", True)], er)
        i = codegen._chain([(wr, "y", False), (wi, "x", False)], ei)
        tr, ti = env.new_tmp(), env.new_tmp()
        lines.append(f"ufloat {tr} = {r if r else env.lit(0.0)};")
        lines.append(f"ufloat {ti} = {i if i else env.lit(0.0)};")
        wr, wi = tr, ti
    px = "tpx[k]"
    py = "tpy[k]"
    if wr is not None:
        px = _fma_neg(wr, "tx[k]", px)
        py = _fma(wr, "ty[k]", py)
    if wi is not None:
        px = _fma(wi, "ty[k]", px)
        py = _fma(wi, "tx[k]", py)
    lines.append(_loop(f"tpx[k] = {px}; tpy[k] = {py};"))
    out.append("{ " + " ".join(lines) + " }")


def emit_thick_tangent(op, env: Env, out: List[str]):
    """Thick elements are linear at fixed dp: the columns get the same
    matrices, without the dispersion terms (those enter the state only)."""
    if isinstance(op, ir.ThickQuad):
        mx, my = codegen.thick_quad_planes(op)
    else:
        mx, my = codegen.thick_bend_planes(op)
    body = []
    for (u, pu), m in ((("tx[k]", "tpx[k]"), mx), (("ty[k]", "tpy[k]"), my)):
        v11, v12, v21, v22 = (env.val(e) for e in m[:4])
        body.append(f"{{ ufloat a = {u}; {u} = fma(a, {v11}, {pu} * {v12}); "
                    f"{pu} = fma(a, {v21}, {pu} * {v22}); }}")
    out.append(_loop(" ".join(body)))


def emit_with_tangent(op, env: Env, out: List[str], observe=None):
    emit_op(op, env, out, observe)
    emit_tangent(op, env, out)


_SOLVE4 = """
// Solve A d = r in place (4x4, partial pivoting, fully unrolled); d in r.
__device__ static inline void ufo_solve4(ufloat A[4][4], ufloat r[4])
{
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int p = c;
#pragma unroll
        for (int i = c + 1; i < 4; ++i)
            if (fabs(A[i][c]) > fabs(A[p][c])) p = i;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            ufloat t = A[c][j];
            A[c][j] = (p == c) ? t : A[p][j];
            if (p != c) A[p][j] = t;
        }
        ufloat t = r[c];
        r[c] = (p == c) ? t : r[p];
        if (p != c) r[p] = t;
#pragma unroll
        for (int i = c + 1; i < 4; ++i) {
            ufloat f = A[i][c] / A[c][c];
#pragma unroll
            for (int j = c; j < 4; ++j) A[i][j] = fma(-f, A[c][j], A[i][j]);
            r[i] = fma(-f, r[c], r[i]);
        }
    }
#pragma unroll
    for (int c = 3; c >= 0; --c) {
        ufloat s = r[c];
#pragma unroll
        for (int j = c + 1; j < 4; ++j) s = fma(-A[c][j], r[j], s);
        r[c] = s / A[c][c];
    }
}
"""


def jvp_kernel(line, flags, turns, parameters, dp, block=256, name="ufo_jvp"):
    """State and the four identity-seeded tangent columns after ``turns``
    turns; a particle stops after the turn at whose end x is NaN.  Output
    [particles][20]: x, px, y, py, then dx_k, dpx_k, dy_k, dpy_k for k = 0..3
    (column k is the derivative with respect to initial coordinate k)."""
    ops = lower(line, flags, parameters)
    src = codegen._coord_sources(parameters, flags, dp)
    env = codegen._env_for(flags, src, ops)
    body, period = codegen._layout(ops, env, emit=emit_with_tangent)
    P = max(1, len(parameters))
    load = codegen._load_block(env, src, P, "row", "    ")
    one = env.lit(1.0)
    zero = env.lit(0.0)
    code = codegen._PRELUDE.format(ufloat="float" if env.fp32 else "double",
                                   nan='__builtin_nanf("")' if env.fp32 else '__builtin_nan("")')
    code += f"""#define NT 4
extern "C" __global__ void __launch_bounds__({block})
{name}(const ufloat *__restrict__ pool, ufloat *__restrict__ out, int particles)
{{
    {codegen._decls(env)}
    ufloat tx[NT], tpx[NT], ty[NT], tpy[NT];
    const int idx = blockIdx.x * {block} + threadIdx.x;
    if (idx >= particles) return;
    const ufloat *__restrict__ row = pool + (size_t)idx * {P};
{chr(10).join(load)}
#pragma unroll
    for (int k = 0; k < NT; ++k) {{
        tx[k] = (k == 0) ? {one} : {zero}; tpx[k] = (k == 1) ? {one} : {zero};
        ty[k] = (k == 2) ? {one} : {zero}; tpy[k] = (k == 3) ? {one} : {zero};
    }}
    for (int turn = 0; turn < {turns}; ++turn) {{
{chr(10).join('        ' + s_ for s_ in body)}
        if (isnan(x)) break;
    }}
    ufloat *o = out + (size_t)idx * 20;
    o[0] = x; o[1] = px; o[2] = y; o[3] = py;
#pragma unroll
    for (int k = 0; k < NT; ++k) {{
        o[4 + 4 * k] = tx[k]; o[5 + 4 * k] = tpx[k]; o[6 + 4 * k] = ty[k]; o[7 + 4 * k] = tpy[k];
    }}
}}
"""
    return codegen.Kernel(name=name, src=code, kind="jvp", fp32=env.fp32,
                          flops_per_turn=codegen._count(ops, env), elements=codegen.element_count(line),
                          block=block, period=period)


def co_kernel(line, flags, parameters, iterations, tol, block=64, name="ufo_co"):
    """Closed orbit per particle by Newton iteration on the one-turn map at
    dp = 0 (ufo/co.py:175-193 fixes dp = 0, oodppo = 1).  The initial guess is
    the particle's x, px, y, py parameters (0 when not named), as the
    reference seeds its simplex (ufo/co.py:208-213).  Output [particles][4];
    an orbit that is lost or does not reach ``tol`` within ``iterations``
    steps is returned as the evaluated iterate with the smallest residual
    (the guess itself if the first evaluation is lost), with ``converged[p]``
    = 0; ``converged[p]`` = 1 when the residual reached ``tol``.  In FP32 the
    iteration also ends at the first step that does not lower the residual."""
    ops = lower(line, flags, parameters)
    src = codegen._coord_sources(parameters, 0, 0.0)
    src["z"] = ("c", 0.0)
    src["dp"] = ("c", 0.0)
    env = codegen._env_for(flags, src, ops)
    body, period = codegen._layout(ops, env, emit=emit_with_tangent)
    P = max(1, len(parameters))
    load = codegen._load_block(env, src, P, "row", "    ")
    one, zero = env.lit(1.0), env.lit(0.0)
    code = codegen._PRELUDE.format(ufloat="float" if env.fp32 else "double",
                                   nan='__builtin_nanf("")' if env.fp32 else '__builtin_nan("")')
    # Every precision keeps the best evaluated iterate and returns it when
    # the residual never reaches tol (cap or loss); FP32, which cannot reach
    # a 1e-15 residual, also stops at the first step that does not lower it.
    huge = "__builtin_huge_valf()" if env.fp32 else "__builtin_huge_val()"
    keep = ("""
        if (!(rmax < best)) break;
        best = rmax; XB[0] = X[0]; XB[1] = X[1]; XB[2] = X[2]; XB[3] = X[3];""" if env.fp32 else """
        if (rmax < best) { best = rmax; XB[0] = X[0]; XB[1] = X[1]; XB[2] = X[2]; XB[3] = X[3]; }""")
    code += "#define NT 4\n" + _SOLVE4
    code += f"""
extern "C" __global__ void __launch_bounds__({block})
{name}(const ufloat *__restrict__ pool, ufloat *__restrict__ orbits, int *__restrict__ steps,
       int *__restrict__ converged, int particles)
{{
    {codegen._decls(env)}
    ufloat tx[NT], tpx[NT], ty[NT], tpy[NT];
    const int idx = blockIdx.x * {block} + threadIdx.x;
    if (idx >= particles) return;
    const ufloat *__restrict__ row = pool + (size_t)idx * {P};
{chr(10).join(load)}
    ufloat X[4] = {{x, px, y, py}};
    ufloat best = {huge}, XB[4] = {{x, px, y, py}};
    bool conv = false;
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
{chr(10).join('            ' + s_ for s_ in body)}
        }}
        ufloat r[4] = {{x - X[0], px - X[1], y - X[2], py - X[3]}};
        if (isnan(x) || isnan(px) || isnan(y) || isnan(py)) break;
        ufloat rmax = fmax(fmax(fabs(r[0]), fabs(r[1])), fmax(fabs(r[2]), fabs(r[3])));
        if (rmax <= {env.lit(tol)}) {{ conv = true; break; }}{keep}
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
    if (!conv) {{ X[0] = XB[0]; X[1] = XB[1]; X[2] = XB[2]; X[3] = XB[3]; }}
    ufloat *o = orbits + (size_t)idx * 4;
    o[0] = X[0]; o[1] = X[1]; o[2] = X[2]; o[3] = X[3];
    steps[idx] = it;
    converged[idx] = conv ? 1 : 0;
}}
"""
    return codegen.Kernel(name=name, src=code, kind="co", fp32=env.fp32,
                          flops_per_turn=codegen._count(ops, env), elements=codegen.element_count(line),
                          block=block, period=period)
