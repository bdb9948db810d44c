"""ORACLE — test infrastructure only (see refgen.py).

Compiles the C text of ``refgen`` with gcc into a shared object (cached by a
hash of source and flags) and runs it through ctypes on numpy arrays.

Build modes:
  ``strict``  -O2 -ffp-contract=off: IEEE evaluation of the reference's
              expressions, left to right, no contraction.
  ``fast``    -O2 -ffast-math -fno-finite-math-only -ffp-contract=fast
              -march=native: the closest CPU analogue of the reference's
              default ``-cl-fast-relaxed-math -cl-mad-enable``, keeping NaN
              tests (a finite-math build may fold ``isnan`` away).
  ``bench_fast``, ``bench_strict``  the CPU-baseline builds: ``fast`` and
              ``strict`` at -O3 -march=native (timing only; the parity
              checks use the two modes above).
FP32 kernels add ``-fsingle-precision-constant`` (the reference's
``-cl-single-precision-constant``, ufo/track.py:56).
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess
import tempfile

import numpy as np

from . import refgen

_CACHE = os.environ.get("UFO_ORACLE_CACHE") or os.path.join(
    tempfile.gettempdir(), f"ufo_oracle_{os.getuid()}")

_FLAGS = {
    "strict": ["-O2", "-ffp-contract=off", "-fno-fast-math"],
    "fast": ["-O2", "-ffast-math", "-fno-finite-math-only", "-ffp-contract=fast", "-march=native"],
    "bench_fast": ["-O3", "-march=native", "-ffast-math", "-fno-finite-math-only", "-ffp-contract=fast"],
    "bench_strict": ["-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math"],
}

FLAGS = _FLAGS


def compile_c(src: str, mode: str = "strict", fp32: bool = False) -> ctypes.CDLL:
    """Build the C text with the mode's flags and load it.  Compiled and
    linked in two steps: gcc adds crtfastmath.o to a -ffast-math link, whose
    constructor switches the loading process to flush-to-zero and
    denormals-are-zero, which would make every later strict run and numpy
    in the same process flush subnormals; the fast modes keep their
    fast-math code generation without it."""
    flags = ["-std=gnu11", "-fPIC", "-fopenmp", "-w"] + _FLAGS[mode]
    if fp32:
        flags.append("-fsingle-precision-constant")
    key = hashlib.sha1((src + "\0" + " ".join(flags) + "\0two-step").encode()).hexdigest()[:20]
    os.makedirs(_CACHE, exist_ok=True)
    so = os.path.join(_CACHE, f"k_{key}.so")
    if not os.path.exists(so):
        c = os.path.join(_CACHE, f"k_{key}.c")
        with open(c, "w") as f:
            f.write(src)
        obj = so + f".{os.getpid()}.o"
        tmp = so + f".{os.getpid()}.tmp"
        subprocess.run(["gcc", *flags, "-c", "-o", obj, c], check=True)
        subprocess.run(["gcc", "-shared", "-fopenmp", "-o", tmp, obj, "-lm"], check=True)
        os.remove(obj)
        os.replace(tmp, so)
    return ctypes.CDLL(so)


def _fp32(flags):
    return not (flags & refgen.FLAG_DOUBLE)


def _dtype(flags):
    return np.float32 if _fp32(flags) else np.float64


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _pool(pool, particles, P, dtype):
    if pool is None:
        pool = np.zeros((particles, max(1, P)), dtype=dtype)
    pool = np.ascontiguousarray(pool, dtype=dtype)
    if pool.ndim == 1:
        pool = pool.reshape(-1, 1)
    if pool.shape[1] == 0:
        pool = np.zeros((pool.shape[0], 1), dtype=dtype)
    return pool


def track(line, pool, flags=0, turns=1, parameters=(), where=(), dp=0., mode="strict"):
    """Reference ``Track`` (ufo/track.py); returns tracks[particles,count,6]
    (entries never written by the kernel stay NaN)."""
    src, count, _ = refgen.track_source(line, flags, turns, parameters, where, dp)
    lib = compile_c(src, mode, _fp32(flags))
    dt = _dtype(flags)
    pool = _pool(pool, None if pool is None else len(pool), len(parameters), dt)
    n = pool.shape[0]
    tracks = np.full((n, count, 6), np.nan, dtype=dt)
    lib.run(_ptr(pool), ctypes.c_int(pool.shape[1]), ctypes.c_int(n), _ptr(tracks))
    return tracks


def stable_aperture(line, pool, flags=0, turns=1000, parameters=(), dp=0., mode="strict"):
    """Reference ``StableAperture`` (ufo/sa.py); returns lost[particles]."""
    src = refgen.sa_source(line, flags, turns, parameters, dp)
    lib = compile_c(src, mode, _fp32(flags))
    dt = _dtype(flags)
    pool = _pool(pool, None, len(parameters), dt)
    n = pool.shape[0]
    lost = np.zeros(n, dtype=np.uint32)
    lib.run(_ptr(pool), ctypes.c_int(pool.shape[1]), ctypes.c_int(n), _ptr(lost))
    return lost


def closed_orbit(line, pool, flags=0, parameters=(), iterations=200, mode="strict"):
    """Reference ``ClosedOrbit`` Nelder-Mead (ufo/co.py); returns
    orbits[particles,4]."""
    src = refgen.co_source(line, flags, parameters, iterations)
    lib = compile_c(src, mode, _fp32(flags))
    dt = _dtype(flags)
    pool = _pool(pool, None, len(parameters), dt)
    n = pool.shape[0]
    orbits = np.zeros((n, 4), dtype=dt)
    lib.run(_ptr(pool), ctypes.c_int(pool.shape[1]), ctypes.c_int(n), _ptr(orbits))
    return orbits


class Optics:
    """Reference ``Optics`` (ufo/optics.py): three-ray linear map kernel plus
    the host Twiss algebra of ufo/optics.py:156-238, step for step."""

    def __init__(self, line, where=(), periodic=True, parameters=(),
                 flags=refgen.FLAG_LINEAR | refgen.FLAG_ACHROMATIC, mode="strict"):
        self.periodic = periodic
        src, self.count, where = refgen.optics_source(line, where, periodic, parameters, flags)
        flat = line.flatten()
        self.where = [(idx, flat[int(idx)]) for idx in where]    # ufo/optics.py:82
        self.lib = compile_c(src, mode, _fp32(flags))
        self.dtype = _dtype(flags)
        self.parameters = np.zeros((1, max(1, len(parameters))), dtype=self.dtype)
        self.ax0 = self.bx0 = self.dx0 = self.dpx0 = 0.
        self.ay0 = self.by0 = self.dy0 = self.dpy0 = 0.

    def run(self):
        track = np.full((3, self.count, 6), np.nan, dtype=self.dtype)
        pool = np.ascontiguousarray(self.parameters, dtype=self.dtype)
        self.lib.run(_ptr(pool), ctypes.c_int(pool.shape[1]), ctypes.c_int(1), _ptr(track))
        self.track = track
        n = self.count
        self.ax, self.ay, self.bx, self.by = (np.empty(n) for _ in range(4))
        self.dx, self.dy, self.dpx, self.dpy = (np.empty(n) for _ in range(4))
        self.mux, self.muy = np.empty(n), np.empty(n)
        if self.periodic:
            re, im, dp = track[0][-1], track[1][-1], track[2][-1]
            cos_nu = (re[0] + im[1]) * 0.5
            if np.absolute(cos_nu) > 1.:
                self.ax0 = self.bx0 = self.qx = np.nan
                return
            self.bx0 = np.absolute(im[0] / np.sin(np.arccos(cos_nu)))
            sin_nu = im[0] / self.bx0
            nu = np.arctan2(sin_nu, cos_nu)
            if nu < 0.:
                nu += 2. * np.pi
            self.ax0 = (re[0] - im[1]) / (2. * np.sin(nu))
            self.qx = nu / (2. * np.pi)
            k = 2. - re[0] - im[1]
            self.dx0 = ((1. - im[1]) * dp[0] + im[0] * dp[1]) / k
            self.dpx0 = ((1. - re[0]) * dp[1] + re[1] * dp[0]) / k
            cos_nu = (re[2] + im[3]) * 0.5
            if np.absolute(cos_nu) > 1.:
                self.ay0 = self.by0 = self.qy = np.nan
                return
            self.by0 = np.absolute(im[2] / np.sin(np.arccos(cos_nu)))
            sin_nu = im[2] / self.by0
            nu = np.arctan2(sin_nu, cos_nu)
            if nu < 0.:
                nu += 2. * np.pi
            self.ay0 = (re[2] - im[3]) / (2. * np.sin(nu))
            self.qy = nu / (2. * np.pi)
        for i, (re, im, dp) in enumerate(zip(track[0], track[1], track[2])):
            g0 = (1. + self.ax0 * self.ax0) / self.bx0
            self.ax[i] = -re[0] * re[1] * self.bx0 + (re[0] * im[1] + im[0] * re[1]) * self.ax0 - im[0] * im[1] * g0
            self.bx[i] = re[0] * re[0] * self.bx0 - re[0] * im[0] * self.ax0 * 2. + im[0] * im[0] * g0
            self.dx[i] = re[0] * self.dx0 + im[0] * self.dpx0 + dp[0]
            self.dpx[i] = re[1] * self.dx0 + im[1] * self.dpx0 + dp[1]
            s = im[0] / np.sqrt(self.bx0 * self.bx[i])
            c = re[0] * np.sqrt(self.bx0 / self.bx[i]) - self.ax0 * s
            self.mux[i] = np.arctan2(s, c)
            g0 = (1. + self.ay0 * self.ay0) / self.by0
            self.ay[i] = -re[2] * re[3] * self.by0 + (re[2] * im[3] + im[2] * re[3]) * self.ay0 - im[2] * im[3] * g0
            self.by[i] = re[2] * re[2] * self.by0 - re[2] * im[2] * self.ay0 * 2. + im[2] * im[2] * g0
            self.dy[i] = re[2] * self.dy0 + im[2] * self.dpy0 + dp[2]
            self.dpy[i] = re[3] * self.dy0 + im[3] * self.dpy0 + dp[3]
            s = im[2] / np.sqrt(self.by0 * self.by[i])
            c = re[2] * np.sqrt(self.by0 / self.by[i]) - self.ay0 * s
            self.muy[i] = np.arctan2(s, c)
