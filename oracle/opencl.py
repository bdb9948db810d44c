"""ORACLE — test infrastructure only (see refgen.py): the reference's own
OpenCL kernels on the GPU's OpenCL device.

The reference builds OpenCL-C kernels from the lattice text and runs them
through pyopencl (ufo/sa.py:94-145, ufo/track.py:84-139).  pyopencl is not
installed, but the ROCm OpenCL runtime is, and on an MI355X box it lists the
GPU as an OpenCL 2.0 device with cl_khr_fp64.  This helper assembles the same
kernels — the element statements restated by ``oracle.refgen`` (the
reference's templates with their literal text), wrapped in the kernel bodies
of ufo/sa.py and ufo/track.py (persistent work-items claiming particles with
``atomic_inc(pool_idx)``) — and builds them with the device's OpenCL compiler
and the reference's default options ``-cl-fast-relaxed-math -cl-mad-enable``
(ufo/__init__.py:22; FP32 adds ``-cl-single-precision-constant``,
ufo/track.py:53), through a small ctypes binding of libOpenCL.

Only ``tests/`` and the ``cpu_baseline`` leg of ``bench.py`` (the reference's
kernel timed on the same GPU beside the CPU baseline) use it; the product
never loads it.  Differences from the reference's wrappers, values unchanged: the pool is a
flat array indexed ``pool + idx * P`` (the reference declares
``pool[][PARAMETERS_COUNT]``) and the tracks buffer ``tracks + (idx * count +
track_idx) * 6`` (``tracks[][count][6]``); ``threads`` must not exceed the
particle count (the reference reads ``pool[get_global_id(0)]`` unchecked).
"""
from __future__ import annotations

import ctypes
import os
import re
import time

import numpy as np

from . import refgen

_CL_DEVICE_TYPE_GPU = 1 << 2
_CL_QUEUE_PROFILING_ENABLE = 1 << 1
_CL_MEM_READ_WRITE = 1 << 0
_CL_MEM_COPY_HOST_PTR = 1 << 5
_CL_PROGRAM_BUILD_LOG = 0x1183
_CL_PROFILING_COMMAND_START = 0x1282
_CL_PROFILING_COMMAND_END = 0x1283
_CL_DEVICE_NAME = 0x102B

DEFAULT_CL_OPTIONS = "-cl-fast-relaxed-math -cl-mad-enable"

_vp = ctypes.c_void_p
_lib = None


class OpenCLError(RuntimeError):
    pass


def _cl():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL("libOpenCL.so.1")
        sig = {
            "clGetPlatformIDs": (ctypes.c_int, [ctypes.c_uint, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_uint)]),
            "clGetDeviceIDs": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_uint, ctypes.POINTER(_vp),
                                              ctypes.POINTER(ctypes.c_uint)]),
            "clGetDeviceInfo": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_size_t, _vp,
                                               ctypes.POINTER(ctypes.c_size_t)]),
            "clCreateContext": (_vp, [_vp, ctypes.c_uint, ctypes.POINTER(_vp), _vp, _vp,
                                      ctypes.POINTER(ctypes.c_int)]),
            "clCreateCommandQueue": (_vp, [_vp, _vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
            "clCreateProgramWithSource": (_vp, [_vp, ctypes.c_uint, ctypes.POINTER(ctypes.c_char_p),
                                                ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int)]),
            "clBuildProgram": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.POINTER(_vp), ctypes.c_char_p, _vp, _vp]),
            "clGetProgramBuildInfo": (ctypes.c_int, [_vp, _vp, ctypes.c_uint, ctypes.c_size_t, _vp,
                                                     ctypes.POINTER(ctypes.c_size_t)]),
            "clCreateKernel": (_vp, [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
            "clCreateBuffer": (_vp, [_vp, ctypes.c_uint64, ctypes.c_size_t, _vp, ctypes.POINTER(ctypes.c_int)]),
            "clSetKernelArg": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_size_t, _vp]),
            "clEnqueueNDRangeKernel": (ctypes.c_int, [_vp, _vp, ctypes.c_uint, _vp, ctypes.POINTER(ctypes.c_size_t),
                                                      _vp, ctypes.c_uint, _vp, ctypes.POINTER(_vp)]),
            "clEnqueueReadBuffer": (ctypes.c_int, [_vp, _vp, ctypes.c_uint, ctypes.c_size_t, ctypes.c_size_t, _vp,
                                                   ctypes.c_uint, _vp, _vp]),
            "clEnqueueWriteBuffer": (ctypes.c_int, [_vp, _vp, ctypes.c_uint, ctypes.c_size_t, ctypes.c_size_t, _vp,
                                                    ctypes.c_uint, _vp, _vp]),
            "clFinish": (ctypes.c_int, [_vp]),
            "clWaitForEvents": (ctypes.c_int, [ctypes.c_uint, ctypes.POINTER(_vp)]),
            "clGetEventProfilingInfo": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_size_t, _vp,
                                                       ctypes.POINTER(ctypes.c_size_t)]),
            "clReleaseEvent": (ctypes.c_int, [_vp]),
            "clReleaseMemObject": (ctypes.c_int, [_vp]),
            "clReleaseKernel": (ctypes.c_int, [_vp]),
            "clReleaseProgram": (ctypes.c_int, [_vp]),
            "clReleaseCommandQueue": (ctypes.c_int, [_vp]),
            "clReleaseContext": (ctypes.c_int, [_vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        _lib = lib
    return _lib


def _check(rc, what):
    if rc != 0:
        raise OpenCLError(f"{what} failed: OpenCL error {rc}")


class Device:
    """Context and profiling queue on the first OpenCL GPU device."""

    def __init__(self):
        cl = _cl()
        n = ctypes.c_uint()
        _check(cl.clGetPlatformIDs(0, None, ctypes.byref(n)), "clGetPlatformIDs")
        plats = (_vp * max(1, n.value))()
        _check(cl.clGetPlatformIDs(n.value, plats, None), "clGetPlatformIDs")
        self.device = None
        for p in plats[:n.value]:
            dev, nd = _vp(), ctypes.c_uint()
            if cl.clGetDeviceIDs(p, _CL_DEVICE_TYPE_GPU, 1, ctypes.byref(dev), ctypes.byref(nd)) == 0 and nd.value:
                self.device = dev
                break
        if self.device is None:
            raise OpenCLError("no OpenCL GPU device")
        buf = ctypes.create_string_buffer(256)
        cl.clGetDeviceInfo(self.device, _CL_DEVICE_NAME, 256, ctypes.cast(buf, _vp), None)
        self.name = buf.value.decode(errors="replace")
        err = ctypes.c_int()
        self.ctx = cl.clCreateContext(None, 1, ctypes.byref(self.device), None, None, ctypes.byref(err))
        _check(err.value, "clCreateContext")
        self.queue = cl.clCreateCommandQueue(self.ctx, self.device, _CL_QUEUE_PROFILING_ENABLE, ctypes.byref(err))
        _check(err.value, "clCreateCommandQueue")

    def build(self, src, options):
        """Program and its kernel ``run``; raises with the build log."""
        cl = _cl()
        err = ctypes.c_int()
        text = ctypes.c_char_p(src.encode())
        prog = cl.clCreateProgramWithSource(self.ctx, 1, ctypes.byref(text), None, ctypes.byref(err))
        _check(err.value, "clCreateProgramWithSource")
        rc = cl.clBuildProgram(prog, 1, ctypes.byref(self.device), options.encode(), None, None)
        if rc != 0:
            size = ctypes.c_size_t()
            cl.clGetProgramBuildInfo(prog, self.device, _CL_PROGRAM_BUILD_LOG, 0, None, ctypes.byref(size))
            log = ctypes.create_string_buffer(size.value + 1)
            cl.clGetProgramBuildInfo(prog, self.device, _CL_PROGRAM_BUILD_LOG, size.value, ctypes.cast(log, _vp),
                                     None)
            cl.clReleaseProgram(prog)
            raise OpenCLError(f"clBuildProgram failed ({rc}):\n{log.value.decode(errors='replace')}")
        kern = cl.clCreateKernel(prog, b"run", ctypes.byref(err))
        _check(err.value, "clCreateKernel")
        return prog, kern

    def buffer(self, array=None, nbytes=None):
        cl = _cl()
        err = ctypes.c_int()
        if array is not None:
            b = cl.clCreateBuffer(self.ctx, _CL_MEM_READ_WRITE | _CL_MEM_COPY_HOST_PTR, array.nbytes,
                                  array.ctypes.data_as(_vp), ctypes.byref(err))
        else:
            b = cl.clCreateBuffer(self.ctx, _CL_MEM_READ_WRITE, max(1, nbytes), None, ctypes.byref(err))
        _check(err.value, "clCreateBuffer")
        return b

    def run(self, kern, buffers, threads):
        """Launch ``run(buffers...)`` over ``threads`` work-items (local size
        left to the runtime, as pyopencl's ``None``); returns the kernel
        time in ms from the event's profiling counters, or the host's wall
        time of enqueue and wait where those read zero (under a profiler)."""
        cl = _cl()
        holders = [_vp(b) for b in buffers]
        for i, h in enumerate(holders):
            _check(cl.clSetKernelArg(kern, i, ctypes.sizeof(_vp), ctypes.byref(h)), "clSetKernelArg")
        gsize = (ctypes.c_size_t * 1)(threads)
        ev = _vp()
        _check(cl.clFinish(self.queue), "clFinish")
        wall = time.perf_counter()
        _check(cl.clEnqueueNDRangeKernel(self.queue, kern, 1, None, gsize, None, 0, None, ctypes.byref(ev)),
               "clEnqueueNDRangeKernel")
        _check(cl.clWaitForEvents(1, ctypes.byref(ev)), "clWaitForEvents")
        wall = time.perf_counter() - wall
        t0, t1 = ctypes.c_uint64(), ctypes.c_uint64()
        cl.clGetEventProfilingInfo(ev, _CL_PROFILING_COMMAND_START, 8, ctypes.byref(t0), None)
        cl.clGetEventProfilingInfo(ev, _CL_PROFILING_COMMAND_END, 8, ctypes.byref(t1), None)
        cl.clReleaseEvent(ev)
        return (t1.value - t0.value) * 1e-6 if t1.value > t0.value else wall * 1e3

    def write(self, buf, array):
        _check(_cl().clEnqueueWriteBuffer(self.queue, buf, 1, 0, array.nbytes, array.ctypes.data_as(_vp), 0, None,
                                          None), "clEnqueueWriteBuffer")

    def read(self, buf, array):
        _check(_cl().clEnqueueReadBuffer(self.queue, buf, 1, 0, array.nbytes, array.ctypes.data_as(_vp), 0, None,
                                         None), "clEnqueueReadBuffer")

    def release(self, *objs):
        cl = _cl()
        for kind, o in objs:
            getattr(cl, {"mem": "clReleaseMemObject", "kernel": "clReleaseKernel",
                         "program": "clReleaseProgram"}[kind])(o)


_device = None


def device():
    """The OpenCL GPU device.  HIP is initialised first: on these boxes a
    process that starts the OpenCL runtime before HIP finds no HIP device
    afterwards."""
    global _device
    if _device is None:
        import torch
        torch.cuda.init()
        _device = Device()
    return _device


def _head(flags):
    ufloat = "double" if flags & refgen.FLAG_DOUBLE else "float"
    text = "#pragma OPENCL EXTENSION cl_khr_fp64 : enable\n" if flags & refgen.FLAG_DOUBLE else ""
    return text + f"#define ufloat {ufloat}\n"


def _init(parameters, fived, dp, indent="        "):
    """Coordinate initialisation of ufo/sa.py:101-111 / ufo/track.py:94-103."""
    text = ""
    for p in ["x", "px", "y", "py", "z", "dp"]:
        src = f"parameters[{list(parameters).index(p)}]" if p in parameters else "0"
        text += f"{indent}{p}  = {src};\n"
    if fived:
        text += f"{indent}z  = 0.;\n{indent}dp = {dp};\n"
    return text + f"{indent}oodppo = 1. / (dp + 1.);\n"


def sa_source(line, flags, turns, parameters, particles, dp=0.0, final_state=False):
    """The kernel of ufo/sa.py:94-126 over the statements of the line.  With
    ``final_state`` it also stores x, px, y, py after its last turn into a
    fourth buffer ``state[particles][4]`` (not part of the reference: the
    checker uses it to tell regular orbits from chaotic ones)."""
    body = "".join(refgen.line_text(line, flags=flags, parameters=parameters))
    P = max(1, len(parameters))
    src = _head(flags)
    extra = ", __global ufloat *state" if final_state else ""
    store = ("        state[(size_t)idx * 4] = x; state[(size_t)idx * 4 + 1] = px;\n"
             "        state[(size_t)idx * 4 + 2] = y; state[(size_t)idx * 4 + 3] = py;\n") if final_state else ""
    src += ("__kernel void run(__global ufloat *pool, __global int *pool_idx, __global uint *lost" + extra + ") {\n"
            "    __global ufloat *parameters;\n"
            "    ufloat x, y, z, px, py, dp;\n"
            "    ufloat oodppo;\n"
            "    int turn, idx = get_global_id(0);\n"
            "    while (1) {\n"
            f"        parameters = pool + (size_t)idx * {P};\n"
            "        lost[idx] = 0;\n")
    src += _init(parameters, flags & refgen.FLAG_FIVED, dp)
    src += (f"        for (turn = 0; turn < {turns}; turn++) {{\n"
            + body +
            "            if (fabs(x) > 1. || fabs(y) > 1. || isnan(x))\n"
            "                break;\n"
            "        }\n"
            "        lost[idx] = turn;\n"
            + store +
            "        idx = atomic_inc(pool_idx);\n"
            f"        if (idx >= {particles})\n"
            "            break;\n"
            "    }\n"
            "}\n")
    return src


def track_source(line, flags, turns, parameters, where, particles, dp=0.0):
    """The kernel of ufo/track.py:56-114 (samples at the fracture points of
    ``where`` and at the end of the line for -1).  Returns (source, count)
    with count = len(where) * turns samples per particle."""
    where = list(where)
    fived = flags & refgen.FLAG_FIVED
    count = len(where) * turns
    dump = "".join(f"tr[{k}] = {c};\n" for k, c in enumerate(["x", "px", "y", "py"] + ([] if fived else ["z", "dp"])))
    dump = ("{ __global ufloat *tr = tracks + ((size_t)idx * " + str(count) + " + track_idx) * 6;\n" + dump
            + "}\ntrack_idx++;\n\n")
    pieces = refgen.line_text(line, fracture=where, flags=flags, parameters=parameters)
    if len(pieces) - 1 + (1 if -1 in where else 0) > len(where):
        raise ValueError("more samples per turn than points in 'where': the tracks rows would overflow")
    code = dump.join(pieces)
    if -1 in where:
        code += dump
    P = max(1, len(parameters))
    src = _head(flags)
    src += ("__kernel void run(__global ufloat *pool, __global int *pool_idx, __global ufloat *tracks) {\n"
            "    __global ufloat *parameters;\n"
            "    ufloat x, y, z, px, py, dp;\n"
            "    ufloat oodppo;\n"
            "    int idx = get_global_id(0);\n"
            "    int track_idx;\n"
            "    while (1) {\n"
            f"        parameters = pool + (size_t)idx * {P};\n")
    src += _init(parameters, fived, dp)
    src += ("        track_idx = 0;\n"
            f"        for (int turn = 0; turn < {turns}; turn++) {{\n"
            + code +
            "            if (isnan(x))\n"
            "                break;\n"
            "        }\n"
            "        idx = atomic_inc(pool_idx);\n"
            f"        if (idx >= {particles})\n"
            "            break;\n"
            "    }\n"
            "}\n")
    return src, count


def _options(flags, options):
    opts = DEFAULT_CL_OPTIONS if options is None else options
    if not flags & refgen.FLAG_DOUBLE:
        opts += " -cl-single-precision-constant"
    return opts


_UNIT_FACTOR = re.compile(r" \* 1\.0(?![0-9])")


def contracted(src):
    """The kernel text with every ``* 1.0`` factor removed (the fraction of a
    whole, unfractured element, ufo/elements.py:102-105): exact in IEEE
    arithmetic, but ``x += px * L * 1.0`` is then ``x += px * L``, which
    ``-cl-mad-enable`` lets the compiler fuse into one multiply-add (the
    ROCm OpenCL compiler does not fuse it while the factor is in the text,
    and does fuse the teapot drifts' ``x += px * (expr)``).  A build of the
    reference's kernel that its own options permit, used as one more sample
    of the reference's arithmetic."""
    return _UNIT_FACTOR.sub("", src)


def stable_aperture(line, pool, flags, turns, parameters, dp=0.0, threads=None, options=None, reps=1,
                    final_state=False, contract=False):
    """Loss turns of the reference's StableAperture kernel on the OpenCL GPU
    (ufo/sa.py:128-145: pool_idx = threads, launch ``(threads,)``).
    Returns (lost, kernel ms of each of ``reps`` launches), and with
    ``final_state`` also x, px, y, py after each particle's last turn
    [particles, 4]."""
    dev = device()
    dt = np.float64 if flags & refgen.FLAG_DOUBLE else np.float32
    pool = np.ascontiguousarray(np.asarray(pool, dtype=dt).reshape(len(pool), -1))
    n = len(pool)
    threads = n if threads is None else int(threads)
    if not 1 <= threads <= n:
        raise ValueError(f"threads = {threads} must be in [1, {n}] (the kernel reads pool[get_global_id(0)])")
    src = sa_source(line, flags, turns, parameters, n, dp, final_state)
    prog, kern = dev.build(contracted(src) if contract else src, _options(flags, options))
    b_pool = dev.buffer(pool)
    b_idx = dev.buffer(nbytes=4)
    b_lost = dev.buffer(nbytes=4 * n)
    bufs = [b_pool, b_idx, b_lost]
    if final_state:
        state = np.full((n, 4), np.nan, dtype=dt)
        bufs.append(dev.buffer(state))
    ms = []
    for _ in range(reps):
        dev.write(b_idx, np.array([threads], dtype=np.int32))
        ms.append(dev.run(kern, bufs, threads))
    lost = np.empty(n, dtype=np.uint32)
    dev.read(b_lost, lost)
    if final_state:
        dev.read(bufs[3], state)
    dev.release(*[("mem", b) for b in bufs], ("kernel", kern), ("program", prog))
    return (lost, ms, state) if final_state else (lost, ms)


def track(line, pool, flags, turns, parameters, where, dp=0.0, threads=None, options=None, reps=1):
    """Samples of the reference's Track kernel on the OpenCL GPU,
    [particles, len(where) * turns, 6]; entries the kernel never writes (after
    a loss, z and dp under FIVED) stay NaN.  Returns (tracks, kernel ms of
    each of ``reps`` launches)."""
    dev = device()
    dt = np.float64 if flags & refgen.FLAG_DOUBLE else np.float32
    pool = np.ascontiguousarray(np.asarray(pool, dtype=dt).reshape(len(pool), -1))
    n = len(pool)
    threads = n if threads is None else int(threads)
    if not 1 <= threads <= n:
        raise ValueError(f"threads = {threads} must be in [1, {n}] (the kernel reads pool[get_global_id(0)])")
    src, count = track_source(line, flags, turns, parameters, where, n, dp)
    prog, kern = dev.build(src, _options(flags, options))
    out = np.full((n, count, 6), np.nan, dtype=dt)
    b_pool = dev.buffer(pool)
    b_idx = dev.buffer(nbytes=4)
    b_tracks = dev.buffer(out)
    ms = []
    for _ in range(reps):
        dev.write(b_idx, np.array([threads], dtype=np.int32))
        dev.write(b_tracks, out)
        ms.append(dev.run(kern, [b_pool, b_idx, b_tracks], threads))
    dev.read(b_tracks, out)
    dev.release(("mem", b_pool), ("mem", b_idx), ("mem", b_tracks), ("kernel", kern), ("program", prog))
    return out, ms


def available():
    """Whether an OpenCL GPU device is present (False without libOpenCL)."""
    if not os.path.exists("/etc/OpenCL/vendors"):
        return False
    try:
        device()
        return True
    except (OSError, OpenCLError):
        return False


def strengths_moved(line):
    """A copy of ``line`` with every magnet strength (k1, k2, k3, k2s, k3s
    and the entries of knl, ksl, dkn, dks) moved up by one ulp: a change of
    the lattice at the level where implementations of the same arithmetic
    differ (constant folding, contraction, operation order)."""
    import copy
    moved = copy.deepcopy(line)
    seen = set()
    for e in moved.flatten():
        if id(e) in seen:          # an element object shared by several periods moves once
            continue
        seen.add(id(e))
        for name in ("k1", "k2", "k3", "k2s", "k3s"):
            v = e.__dict__.get(name)
            if isinstance(v, float) and v != 0.0:
                e.__dict__[name] = float(np.nextafter(v, np.inf))
        for name in ("knl", "ksl", "dkn", "dks"):
            v = e.__dict__.get(name)
            if isinstance(v, list):
                e.__dict__[name] = [float(np.nextafter(u, np.inf)) if isinstance(u, float) and u != 0.0 else u
                                    for u in v]
    return moved


def reference_variants(line, pool, flags, turns, parameters, threads, dp=0.0):
    """Loss turns and final states of the reference's StableAperture kernel
    on the OpenCL GPU in six builds of its arithmetic: ``default`` (its
    default options), ``strict`` (no options), ``up`` and ``down`` (default
    options, every input moved by one ulp), ``strengths`` (the lattice with
    every magnet strength moved by one ulp, ``strengths_moved``) and
    ``contracted`` (the ``contracted`` text).  Returns {name: (lost, state,
    kernel ms)}."""
    dt = np.float64 if flags & refgen.FLAG_DOUBLE else np.float32
    base = np.ascontiguousarray(np.asarray(pool, dtype=dt).reshape(len(pool), -1))
    moved_line = strengths_moved(line)
    builds = {"default": (line, base, None, False), "strict": (line, base, "", False),
              "up": (line, np.nextafter(base, dt(np.inf)), None, False),
              "down": (line, np.nextafter(base, dt(-np.inf)), None, False),
              "strengths": (moved_line, base, None, False), "contracted": (line, base, None, True)}
    out = {}
    for name, (ln, arr, opts, fuse) in builds.items():
        lost, ms, state = stable_aperture(ln, arr, flags, turns, parameters, dp=dp, threads=threads, options=opts,
                                          final_state=True, contract=fuse)
        out[name] = (lost.astype(np.int64), state.astype(np.float64), ms[0])
    return out


def regular_mask(variants, turns, tol=1e-9, leave_out=()):
    """Particles whose loss turn is regular over ``variants`` (except those
    named in ``leave_out``): every build gives the default build's loss turn
    and, for survivors, a final state within ``tol`` of it (relative, floor
    the rms of each coordinate over the survivors): over long runs a chaotic
    particle near the edge of the dynamic aperture can survive in every run
    by chance while its runs have long parted."""
    ref_lost, ref_state, _ = variants["default"]
    alive = ref_lost == turns
    rms = np.sqrt(np.mean(ref_state[alive] ** 2, axis=0)) if alive.any() else np.ones(ref_state.shape[1])
    rms = np.where(rms > 0, rms, 1.0)
    scale = np.maximum(np.abs(np.nan_to_num(ref_state, nan=0.0, posinf=0.0, neginf=0.0)), rms)
    regular = np.ones(len(ref_lost), dtype=bool)
    for name, (lost, state, _) in variants.items():
        if name == "default" or name in leave_out:
            continue
        regular &= lost == ref_lost
        with np.errstate(invalid="ignore", over="ignore"):
            d = np.abs(np.nan_to_num(state - ref_state, nan=0.0)) / scale
        regular &= ~alive | (np.max(d, axis=1) <= tol)
    return regular
