"""Test helper: the emitted Track kernel compiled for the host.

The HIP source that ``ufo_amd.codegen.track_kernel`` emits is plain C++
arithmetic (``fma``, ``+``, ``*``, comparisons); compiled by g++ with
``-ffp-contract=off`` it performs the same IEEE operations in the same order
as the GPU, so on lattices without transcendental functions (teapot KICK
paths) it reproduces the GPU result bit for bit.  The GPU tests use it to
check that the device executes exactly what was emitted; the parity bar
itself is against the oracle.  Never used by the product.
"""
import ctypes
import hashlib
import os
import subprocess
import tempfile

import numpy as np

_PRELUDE = r"""
#include <cmath>
#include <cstddef>
using std::fma; using std::sqrt; using std::fabs; using std::sin; using std::cos; using std::tan;
using std::sinh; using std::cosh; using std::isnan; using std::rint; using std::fmod; using std::ldexp;
struct ufo_dim3 { unsigned x, y, z; };
static ufo_dim3 blockIdx, threadIdx;
#define __global__
#define __device__
#define __launch_bounds__(n)
#define UFO_STREAM_STORE(v, p) (*(p) = (v))
"""

_DRIVER = r"""
extern "C" void host_run(const ufloat *pool, ufloat *tracks, ufloat *state, int particles, int block,
                         int turn0, int turn1)
{
    int blocks = (particles + block - 1) / block;
    for (int b = 0; b < blocks; ++b) {
        blockIdx.x = b;
        for (int t = 0; t < block; ++t) {
            threadIdx.x = t;
            %s(pool, tracks, state, particles, turn0, turn1);
        }
    }
}
"""

_CACHE = os.path.join(tempfile.gettempdir(), f"ufo_host_emulation_{os.getuid()}")


def compile_track(kernel):
    src = _PRELUDE + kernel.src + (_DRIVER % kernel.name)
    key = hashlib.sha1(src.encode()).hexdigest()[:20]
    os.makedirs(_CACHE, exist_ok=True)
    so = os.path.join(_CACHE, f"t_{key}.so")
    if not os.path.exists(so):
        cpp = os.path.join(_CACHE, f"t_{key}.cpp")
        with open(cpp, "w") as f:
            f.write(src)
        tmp = so + f".{os.getpid()}"
        subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                        "-w", "-o", tmp, cpp], check=True)
        os.replace(tmp, so)
    return ctypes.CDLL(so)


def run_track(kernel, pool, samples, splits=(), fill=np.nan):
    """Host run of a generated Track kernel: returns [particles, samples, 6].
    ``splits`` are the turns at which the run is cut into launches that
    resume from the saved state, as Track.run does; ``fill`` is the value
    the samples buffer holds before the first launch."""
    lib = compile_track(kernel)
    dt = np.float32 if kernel.fp32 else np.float64
    pool = np.ascontiguousarray(pool, dtype=dt)
    n = pool.shape[0]
    out = np.full((samples, 6, n), fill, dtype=dt)
    state = np.full((4, max(1, n)), np.nan, dtype=dt)
    bounds = [0] + sorted(splits) + [kernel.turns]
    for t0, t1 in zip(bounds[:-1], bounds[1:]):
        lib.host_run(pool.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
                     state.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(n), ctypes.c_int(kernel.block),
                     ctypes.c_int(t0), ctypes.c_int(t1))
    return out.transpose(2, 0, 1)


_SANITIZED_MAIN = r"""
#include <cstdio>
#include <cstdlib>
int main() {
    const int n = %(n)d, P = %(P)d, S = %(S)d;
    // every buffer exactly as large as the launch needs, so that any access
    // past an end is reported
    ufloat *pool = (ufloat *)malloc(sizeof(ufloat) * n * P);
    ufloat *tracks = (ufloat *)malloc(sizeof(ufloat) * S * 6 * n);
    ufloat *state = (ufloat *)malloc(sizeof(ufloat) * 4 * n);
    for (int i = 0; i < n * P; ++i) pool[i] = (ufloat)(%(scale)r * ((i * 7919) %% 101 - 50));
    for (int i = 0; i < S * 6 * n; ++i) tracks[i] = 0;
    const int b[] = {%(bounds)s};
    for (int k = 0; k + 1 < (int)(sizeof b / sizeof b[0]); ++k)
        host_run(pool, tracks, state, n, %(block)d, b[k], b[k + 1]);
    free(pool); free(tracks); free(state);
    return 0;
}
"""


def sanitized_run(kernel, particles, columns, splits=(), scale=1e-4):
    """Build the emitted kernel, the launch loop and exactly sized buffers
    as a host program under AddressSanitizer and UndefinedBehaviorSanitizer
    and run it over the turns cut at ``splits``; pool entries are ``scale``
    times integers in [-50, 50]; returns (exit status, sanitizer report)."""
    samples = kernel.dumps_per_turn * kernel.turns
    bounds = ", ".join(str(t) for t in [0] + sorted(splits) + [kernel.turns])
    src = (_PRELUDE + kernel.src + (_DRIVER % kernel.name)
           + _SANITIZED_MAIN % dict(n=particles, P=max(1, columns), S=samples, block=kernel.block, bounds=bounds,
                                    scale=float(scale)))
    d = tempfile.mkdtemp(prefix="ufo_sanitized_")
    cpp, exe = os.path.join(d, "k.cpp"), os.path.join(d, "k")
    with open(cpp, "w") as f:
        f.write(src)
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-w", "-o", exe, cpp], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stderr


# buffers and call of the other kernel kinds: (allocations, call, frees);
# n = particles (rows of pool), S = samples per ray (Optics)
_KINDS = {
    "optics": ("ufloat *tracks = (ufloat *)malloc(sizeof(ufloat) * 3 * S * 6);", "%s(pool, tracks, S);",
               "free(tracks);"),
    "jvp": ("ufloat *out = (ufloat *)malloc(sizeof(ufloat) * n * 20);", "%s(pool, out, n);", "free(out);"),
    "co": ("ufloat *orbits = (ufloat *)malloc(sizeof(ufloat) * n * 4); int *steps = (int *)malloc(sizeof(int) * n);"
           " int *conv = (int *)malloc(sizeof(int) * n);",
           "%s(pool, orbits, steps, conv, n);", "free(orbits); free(steps); free(conv);"),
}

_KIND_MAIN = r"""
#include <cstdio>
#include <cstdlib>
int main() {
    const int n = %(n)d, P = %(P)d, S = %(S)d, blocks = %(blocks)d, block = %(block)d;
    ufloat *pool = (ufloat *)malloc(sizeof(ufloat) * n * P);
    for (int i = 0; i < n * P; ++i) pool[i] = (ufloat)(1e-5 * ((i * 7919) %% 101 - 50));
    %(alloc)s
    for (int b = 0; b < blocks; ++b) {
        blockIdx.x = b;
        for (int t = 0; t < block; ++t) { threadIdx.x = t; %(call)s }
    }
    %(free)s
    free(pool);
    (void)S;
    return 0;
}
"""


def sanitized_run_kind(kernel, particles, columns, samples=0):
    """``sanitized_run`` for the Optics (one 64-lane block, three rays,
    ``samples`` per ray), JVP and ClosedOrbit kernels: every buffer exactly
    as large as the launch needs; returns (exit status, sanitizer report)."""
    alloc, call, free = _KINDS[kernel.kind]
    n = 1 if kernel.kind == "optics" else particles
    blocks = 1 if kernel.kind == "optics" else (particles + kernel.block - 1) // kernel.block
    src = (_PRELUDE + kernel.src
           + _KIND_MAIN % dict(n=n, P=max(1, columns), S=samples, blocks=blocks, block=kernel.block, alloc=alloc,
                               call=call % kernel.name, free=free))
    d = tempfile.mkdtemp(prefix="ufo_sanitized_")
    cpp, exe = os.path.join(d, "k.cpp"), os.path.join(d, "k")
    with open(cpp, "w") as f:
        f.write(src)
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-w", "-o", exe, cpp], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stderr


_ORBIT_DRIVERS = {
    "co": r"""
extern "C" void host_run(const ufloat *pool, ufloat *orbits, int *steps, int *conv, int n, int block)
{
    for (int b = 0; b < (n + block - 1) / block; ++b) {
        blockIdx.x = b;
        for (int t = 0; t < block; ++t) { threadIdx.x = t; %s(pool, orbits, steps, conv, n); }
    }
}
""",
    "co_seed": r"""
extern "C" void host_run(const ufloat *table, const ufloat *guess, ufloat *orbits, int *steps, int *conv, int n,
                         int block)
{
    for (int b = 0; b < (n + block - 1) / block; ++b) {
        blockIdx.x = b;
        for (int t = 0; t < block; ++t) { threadIdx.x = t; %s(table, guess, orbits, steps, conv, n); }
    }
}
""",
}


def _compile_kind(kernel):
    src = _PRELUDE + kernel.src + (_ORBIT_DRIVERS[kernel.kind] % kernel.name)
    key = hashlib.sha1(src.encode()).hexdigest()[:20]
    os.makedirs(_CACHE, exist_ok=True)
    so = os.path.join(_CACHE, f"o_{key}.so")
    if not os.path.exists(so):
        cpp = os.path.join(_CACHE, f"o_{key}.cpp")
        with open(cpp, "w") as f:
            f.write(src)
        tmp = so + f".{os.getpid()}"
        subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                        "-w", "-o", tmp, cpp], check=True)
        os.replace(tmp, so)
    return ctypes.CDLL(so)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def run_orbits(kernel, pool):
    """Host run of a generated ClosedOrbit kernel (``tangent.co_kernel``)
    over the rows of ``pool``: (orbits [n, 4], steps [n], converged [n])."""
    lib = _compile_kind(kernel)
    dt = np.float32 if kernel.fp32 else np.float64
    pool = np.ascontiguousarray(pool, dtype=dt)
    n = pool.shape[0]
    orbits = np.full((n, 4), np.nan, dtype=dt)
    steps = np.full(n, -1, dtype=np.int32)
    conv = np.full(n, -1, dtype=np.int32)
    lib.host_run(_ptr(pool), _ptr(orbits), _ptr(steps), _ptr(conv), ctypes.c_int(n), ctypes.c_int(kernel.block))
    return orbits, steps, conv != 0


def run_seed_orbits(kernel, table, guess):
    """Host run of a generated per-seed closed-orbit kernel
    (``batch_adjoint.co_seed_kernel``): (orbits, steps, converged)."""
    lib = _compile_kind(kernel)
    table = np.ascontiguousarray(table, dtype=np.float64)
    guess = np.ascontiguousarray(guess, dtype=np.float64)
    n = guess.shape[0]
    orbits = np.full((n, 4), np.nan)
    steps = np.full(n, -1, dtype=np.int32)
    conv = np.full(n, -1, dtype=np.int32)
    lib.host_run(_ptr(table), _ptr(guess), _ptr(orbits), _ptr(steps), _ptr(conv), ctypes.c_int(n),
                 ctypes.c_int(kernel.block))
    return orbits, steps, conv != 0


# ---------------------------------------------------------------------------
# Kernels with workgroup barriers and LDS atomics (the adjoint kernels): one
# host thread per lane, a C++20 barrier per workgroup, std::atomic_ref for
# the FP64 atomic adds.
# ---------------------------------------------------------------------------

_THREADED_PRELUDE = r"""
#include <cmath>
#include <cstddef>
#include <atomic>
#include <barrier>
#include <thread>
#include <vector>
using std::fma; using std::sqrt; using std::fabs; using std::sin; using std::cos; using std::tan;
using std::sinh; using std::cosh; using std::isnan; using std::rint; using std::fmod; using std::ldexp;
typedef double ufloat;   // the adjoint kernels compute in FP64 only
struct ufo_dim3 { unsigned x, y, z; };
static thread_local ufo_dim3 blockIdx, threadIdx;
static std::barrier<> *ufo_workgroup;
#define __global__
#define __device__
#define __shared__ static
#define __launch_bounds__(n)
static inline void __syncthreads() { ufo_workgroup->arrive_and_wait(); }
static inline void unsafeAtomicAdd(double *p, double v) { std::atomic_ref<double>(*p).fetch_add(v); }
template <class F> static void ufo_grid(int grid, int block, F body)
{
    for (int b = 0; b < grid; ++b) {
        std::barrier<> wg(block);
        ufo_workgroup = &wg;
        std::vector<std::thread> lanes;
        for (int t = 0; t < block; ++t)
            lanes.emplace_back([=] { blockIdx = {unsigned(b), 0, 0}; threadIdx = {unsigned(t), 0, 0}; body(); });
        for (auto &l : lanes) l.join();
    }
}
"""

_ADJ_DRIVER = r"""
extern "C" void host_table(int grid, const ufloat *th, ufloat *ct)
{ ufo_grid(grid, 64, [=] { %(table)s(th, ct); }); }
extern "C" void host_fwd(int grid, const ufloat *pool, const ufloat *ct, ufloat *ckpt, ufloat *xout, int *lost, int n)
{ ufo_grid(grid, 256, [=] { %(fwd)s(pool, ct, ckpt, xout, lost, n); }); }
extern "C" void host_bwd(int grid, int block, const ufloat *pool, const ufloat *ct, const ufloat *ckpt,
                         const ufloat *lam, const int *lost, ufloat *gx0, ufloat *gpart, int n)
{ ufo_grid(grid, block, [=] { %(bwd)s(pool, ct, ckpt, lam, lost, gx0, gpart, n); }); }
"""


class HostAdjoint:
    """The kernels of ``adjoint.adjoint_kernels`` built for the host and
    launched as ``AdjointTracker`` launches them: ``forward(x0, theta)`` ->
    (final state [n, 4], lost [n], saved), ``backward(lam, saved)`` ->
    (dL/dx0 [n, 4], dL/dtheta)."""

    def __init__(self, line, flags, turns, coords, lattice_parameters, every=8, dp=0.0):
        from ufo_amd import adjoint
        self.kfwd, self.kbwd, self.ktab, self.layout, self.ncheck = adjoint.adjoint_kernels(
            line, flags, turns, coords, lattice_parameters, dp=dp, every=every)
        self.turns, self.npar = turns, max(1, len(lattice_parameters))
        # one translation unit, each kernel's source in a namespace of its
        # own (the sources repeat their helper functions)
        kernels = (self.ktab, self.kfwd, self.kbwd)
        body = "\n".join(f"namespace k{i} {{\n{k.src}\n}}" for i, k in enumerate(kernels))
        src = _THREADED_PRELUDE + body + _ADJ_DRIVER % dict(table=f"k0::{self.ktab.name}", fwd=f"k1::{self.kfwd.name}",
                                                            bwd=f"k2::{self.kbwd.name}")
        key = hashlib.sha1(src.encode()).hexdigest()[:20]
        os.makedirs(_CACHE, exist_ok=True)
        so = os.path.join(_CACHE, f"a_{key}.so")
        if not os.path.exists(so):
            cpp = os.path.join(_CACHE, f"a_{key}.cpp")
            with open(cpp, "w") as f:
                f.write(src)
            tmp = so + f".{os.getpid()}"
            subprocess.run(["g++", "-std=c++20", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                            "-pthread", "-w", "-o", tmp, cpp], check=True)
            os.replace(tmp, so)
        self.lib = ctypes.CDLL(so)

    def forward(self, x0, theta):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        n = x0.shape[0]
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(-1))
        ct = np.zeros(self.layout["nslots"])
        grid = max(1, (max(self.layout["reps"], 1) + 63) // 64)
        self.lib.host_table(ctypes.c_int(grid), _ptr(th), _ptr(ct))
        ckpt = np.full((self.ncheck, 4, n), np.nan)
        xf = np.full((n, 4), np.nan)
        lost = np.full(n, -1, dtype=np.int32)
        self.lib.host_fwd(ctypes.c_int((n + 255) // 256), _ptr(x0), _ptr(ct), _ptr(ckpt), _ptr(xf), _ptr(lost),
                          ctypes.c_int(n))
        return xf, lost, (x0, ct, ckpt, lost)

    def backward(self, lam, saved):
        x0, ct, ckpt, lost = saved
        n = x0.shape[0]
        lam = np.ascontiguousarray(np.nan_to_num(lam, nan=0.0), dtype=np.float64)
        block = self.kbwd.block
        nb = max(1, (n + block - 1) // block)
        gx0 = np.zeros((n, 4))
        gpart = np.full((nb, self.npar), np.nan)
        self.lib.host_bwd(ctypes.c_int(nb), ctypes.c_int(block), _ptr(x0), _ptr(ct), _ptr(ckpt), _ptr(lam),
                          _ptr(lost), _ptr(gx0), _ptr(gpart), ctypes.c_int(n))
        g = gpart.sum(axis=0)
        scale = np.asarray(self.layout["gscale"], dtype=np.float64)
        g[:scale.size] *= scale     # the factors the kernel leaves to the reduction
        return gx0, g


_BADJ_DRIVER = r"""
extern "C" void host_fwd(int grid, const ufloat *pool, const ufloat *table, const int *ts, const int *t0,
                         const int *tc, ufloat *ckpt, ufloat *xout, int *lost, int n)
{ ufo_grid(grid, 64, [=] { %(fwd)s(pool, table, ts, t0, tc, ckpt, xout, lost, n); }); }
extern "C" void host_bwd(int grid, const ufloat *pool, const ufloat *table, const int *ts, const int *t0,
                         const int *tc, const ufloat *ckpt, const ufloat *lam, const int *lost, ufloat *gx0,
                         ufloat *gseed, int n)
{ ufo_grid(grid, 64, [=] { %(bwd)s(pool, table, ts, t0, tc, ckpt, lam, lost, gx0, gseed, n); }); }
"""


class HostBatchedAdjoint:
    """The two kernels of a ``batch_adjoint.BatchedAdjoint`` (built on the
    CPU; it needs no device until it launches) on the host, launched over
    its 64-particle tiles as it launches them: ``forward(x0)`` -> (final
    state, lost, saved), ``backward(lam, saved)`` -> (dL/dx0, dL/dtheta per
    seed [seeds, D])."""

    def __init__(self, batched):
        self.b = batched
        kernels = (batched.kfwd, batched.kbwd)
        body = "\n".join(f"namespace k{i} {{\n{k.src}\n}}" for i, k in enumerate(kernels))
        src = _THREADED_PRELUDE.replace("#define __shared__ static", "#define __shared__ static\n#define __constant__")
        src += body + _BADJ_DRIVER % dict(fwd=f"k0::{batched.kfwd.name}", bwd=f"k1::{batched.kbwd.name}")
        key = hashlib.sha1(src.encode()).hexdigest()[:20]
        os.makedirs(_CACHE, exist_ok=True)
        so = os.path.join(_CACHE, f"b_{key}.so")
        if not os.path.exists(so):
            cpp = os.path.join(_CACHE, f"b_{key}.cpp")
            with open(cpp, "w") as f:
                f.write(src)
            tmp = so + f".{os.getpid()}"
            subprocess.run(["g++", "-std=c++20", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                            "-pthread", "-w", "-o", tmp, cpp], check=True)
            os.replace(tmp, so)
        self.lib = ctypes.CDLL(so)
        self.table = np.ascontiguousarray(batched.table, dtype=np.float64)
        self.tiles = [np.ascontiguousarray(t, dtype=np.int32) for t in batched._tiles]

    def forward(self, x0):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        n = self.b.particles
        assert x0.shape == (n, len(self.b.coords))
        ckpt = np.full((self.b.ncheck, 4, max(1, n)), np.nan)
        xf = np.full((n, 4), np.nan)
        lost = np.full(n, -1, dtype=np.int32)
        ts, t0, tc = self.tiles
        self.lib.host_fwd(ctypes.c_int(len(ts)), _ptr(x0), _ptr(self.table), _ptr(ts), _ptr(t0), _ptr(tc),
                          _ptr(ckpt), _ptr(xf), _ptr(lost), ctypes.c_int(n))
        return xf, lost, (x0, ckpt, lost)

    def backward(self, lam, saved):
        x0, ckpt, lost = saved
        n = self.b.particles
        lam = np.ascontiguousarray(np.nan_to_num(lam, nan=0.0), dtype=np.float64)
        gx0 = np.zeros((n, 4))
        gseed = np.zeros((len(self.b.counts), self.b.ndiff))
        ts, t0, tc = self.tiles
        self.lib.host_bwd(ctypes.c_int(len(ts)), _ptr(x0), _ptr(self.table), _ptr(ts), _ptr(t0), _ptr(tc),
                          _ptr(ckpt), _ptr(lam), _ptr(lost), _ptr(gx0), _ptr(gseed), ctypes.c_int(n))
        return gx0, gseed
