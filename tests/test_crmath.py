"""The correctly rounded functions of the kernels' per-particle invariants
(ufo_amd/crmath.py), built for the host from the same text the kernels
include, against mpmath at 300 bits; and the oracle's long-double route to
the same functions.  CPU only."""
import ctypes
import os
import subprocess
import tempfile

import mpmath
import numpy as np
import pytest

from ufo_amd import crmath

NAMES = ["sin", "cos", "tan", "sinh", "cosh"]
MP = [mpmath.sin, mpmath.cos, mpmath.tan, mpmath.sinh, mpmath.cosh]


def _build(text, name, lang="c++"):
    d = tempfile.mkdtemp(prefix="ufo_crmath_")
    src = os.path.join(d, name + (".cpp" if lang == "c++" else ".c"))
    so = os.path.join(d, name + ".so")
    open(src, "w").write(text)
    cc = "g++" if lang == "c++" else "gcc"
    subprocess.run([cc, "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-w", "-o", so, src, "-lm"],
                   check=True)
    return ctypes.CDLL(so)


def _exact(f, x):
    with mpmath.workprec(300):
        return np.array([float(f(mpmath.mpf(v))) for v in x])


@pytest.fixture(scope="module")
def crlib():
    text = ("#include <cmath>\nusing std::fma; using std::rint; using std::fmod; using std::ldexp; using std::fabs;\n"
            "using std::sin; using std::cos; using std::tan; using std::sinh; using std::cosh;\n#define __device__\n"
            + crmath.source() +
            '\nextern "C" void run(int f, const double *x, double *y, int n) {\n'
            "  for (int i = 0; i < n; ++i) { double v = x[i];\n"
            "    y[i] = f == 0 ? ufo_cr_sin(v) : f == 1 ? ufo_cr_cos(v) : f == 2 ? ufo_cr_tan(v)"
            " : f == 3 ? ufo_cr_sinh(v) : ufo_cr_cosh(v); } }\n")
    return _build(text, "cr")


@pytest.mark.parametrize("lo,hi", [(-0.8, 0.8), (-4.0, 4.0), (-25.0, 25.0)])
def test_double_double_functions_are_correctly_rounded(crlib, lo, hi):
    """Every result on 3000 random arguments per function and range equals
    the correctly rounded value (the argument reduction by pi/2 and ln 2,
    the series, the quadrant and sign logic all take part)."""
    rng = np.random.default_rng(int(hi * 10))
    for fi, f in enumerate(MP):
        x = rng.uniform(lo, hi, 3000)
        y = np.zeros_like(x)
        crlib.run(fi, x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p), len(x))
        bad = np.flatnonzero(y != _exact(f, x))
        assert len(bad) == 0, f"{NAMES[fi]}: {len(bad)} results not correctly rounded, e.g. x = {x[bad[:3]]}"


def test_special_arguments(crlib):
    """Zero, tiny, quadrant boundaries, NaN and huge arguments."""
    x = np.array([0.0, 1e-300, -1e-20, np.pi / 2, np.pi, -np.pi / 4, 1.0, -1.0, 700.5, np.nan, 1e7])
    for fi, f in enumerate(MP):
        y = np.zeros_like(x)
        crlib.run(fi, x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p), len(x))
        finite = np.isfinite(x) & (np.abs(x) < 700) if fi >= 3 else np.isfinite(x) & (np.abs(x) < 1e6)
        np.testing.assert_array_equal(y[finite], _exact(f, x[finite]))
        assert np.isnan(y[9])


def test_oracle_long_double_route_is_nearly_correctly_rounded():
    """The oracle rounds its platform's long double functions once: at most
    a few in ten thousand results miss the correctly rounded double (the
    host's double sinh and cosh miss about a quarter)."""
    lib = _build("#include <math.h>\nvoid run(int f, const double *x, double *y, int n) {\n"
                 "  for (int i = 0; i < n; ++i) { long double v = x[i];\n"
                 "    y[i] = (double)(f == 0 ? sinl(v) : f == 1 ? cosl(v) : f == 2 ? tanl(v) : f == 3 ? sinhl(v)"
                 " : coshl(v)); } }\n", "ld", lang="c")
    rng = np.random.default_rng(9)
    for fi, f in enumerate(MP):
        x = rng.uniform(-4, 4, 4000)
        y = np.zeros_like(x)
        lib.run(fi, x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p), len(x))
        assert np.sum(y != _exact(f, x)) <= 8, NAMES[fi]
