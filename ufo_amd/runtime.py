"""ctypes binding of libufo_hip.so (include/ufo_hip.h) and the device layer
the drivers use: contexts, hipRTC modules, launches on torch's current
stream.

``list_devices(device=None)`` and ``context(device)`` keep the reference's
names and behaviour (ufo/cl_utils.py:12-34): the first prints ``idx: name``
for every device or returns a descriptor, the second returns a context for
the simulation classes.

torch is imported before the library is loaded so that the library binds to
the HIP runtime torch already loaded (one runtime per process, shared
streams).
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (must precede the library load, see module doc)

from .build import LIB

_lib = None
_lock = threading.Lock()


class UfoError(RuntimeError):
    pass


class _Cfg(ctypes.Structure):
    _fields_ = [("grid", ctypes.c_uint * 3), ("block", ctypes.c_uint * 3), ("shared_bytes", ctypes.c_uint)]


class _Props(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64), ("compute_units", ctypes.c_int),
                ("clock_khz", ctypes.c_int), ("warp_size", ctypes.c_int), ("total_mem", ctypes.c_size_t)]


_SIGS = {
    "ufo_last_error": (ctypes.c_char_p, []),
    "ufo_version": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "ufo_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "ufo_device_name": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
    "ufo_device_props_get": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_Props)]),
    "ufo_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "ufo_context_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ufo_compile": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                   ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t),
                                   ctypes.c_char_p, ctypes.c_size_t]),
    "ufo_code_free": (None, [ctypes.c_void_p]),
    "ufo_module_compile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                          ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_size_t]),
    "ufo_module_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "ufo_module_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ufo_kernel_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int)]),
    "ufo_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(_Cfg),
                                  ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_void_p]),
    "ufo_launch_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(_Cfg),
                                          ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_size_t,
                                          ctypes.c_size_t, ctypes.c_void_p]),
    "ufo_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "ufo_malloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "ufo_free": (ctypes.c_int, [ctypes.c_void_p]),
    "ufo_memcpy_htod": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "ufo_memcpy_dtoh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "ufo_graph_begin": (ctypes.c_int, [ctypes.c_void_p]),
    "ufo_graph_end": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "ufo_graph_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "ufo_graph_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ufo_time_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(_Cfg),
                                       ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
}


def lib():
    """The loaded libufo_hip.so; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB):
                raise UfoError(f"{LIB} is missing: build it with __graft_entry__.build() "
                               "or `python -m ufo_amd.build`")
            handle = ctypes.CDLL(LIB)
            for name, (res, args) in _SIGS.items():
                fn_ = getattr(handle, name)
                fn_.restype = res
                fn_.argtypes = args
            _lib = handle
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().ufo_last_error().decode(errors="replace")
        raise UfoError(f"{what}: {msg}" if what else msg)


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().ufo_device_count(ctypes.byref(n)), "ufo_device_count")
    return n.value


class Device:
    """Descriptor of one HIP device."""

    def __init__(self, index: int):
        p = _Props()
        check(lib().ufo_device_props_get(index, ctypes.byref(p)), "ufo_device_props_get")
        self.index = index
        self.name = p.name.decode()
        self.arch = p.arch.decode().split(":")[0]
        self.compute_units = p.compute_units
        self.clock_khz = p.clock_khz
        self.total_mem = p.total_mem

    @property
    def fp64_peak(self) -> float:
        """FP64 vector peak in FLOP/s: CUs x 128 FLOP/clk x peak clock."""
        return self.compute_units * 128.0 * self.clock_khz * 1e3

    def __repr__(self):
        return f"{self.name} ({self.arch}, {self.compute_units} CUs, {self.clock_khz / 1e3:.0f} MHz)"


def list_devices(device=None):
    """Print ``idx:   name`` for every HIP device, or return the descriptor
    of device ``device`` (ufo/cl_utils.py:12-28)."""
    n = device_count()
    for i in range(n):
        if device == i:
            return Device(i)
        if not device:
            print(f"{i}:   {Device(i).name}")
    return None


class Context:
    """A device plus its compiled-module cache (replaces the OpenCL context
    of ufo/cl_utils.py:30-34)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().ufo_context_create(device, ctypes.byref(h)), f"context on device {device}")
        self.handle = h
        self.device = Device(device)
        self.torch_device = torch.device("cuda", device)
        self._modules = {}

    def module(self, src: str, options: str = "") -> "Module":
        key = (src, options)
        m = self._modules.get(key)
        if m is None:
            m = Module(self, src, options)
            self._modules[key] = m
        return m

    def __del__(self):
        try:
            if _lib is not None and self.handle:
                _lib.ufo_context_destroy(self.handle)
        except Exception:
            pass


def context(device=0) -> Context:
    """A context for the simulation classes (ufo/cl_utils.py:30-34)."""
    return Context(device)


_default = None


def default_context() -> Context:
    global _default
    if _default is None:
        if device_count() < 1:
            raise UfoError("no HIP device: the tracking kernels need an MI355X (gfx950) GPU")
        _default = Context(0)
    return _default


def compile_only(src: str, options: str = "", arch: str = "gfx950") -> bytes:
    """hipRTC build without a device (used to check generated sources)."""
    code = ctypes.c_void_p()
    size = ctypes.c_size_t()
    log = ctypes.create_string_buffer(1 << 16)
    rc = lib().ufo_compile(src.encode(), options.encode(), arch.encode(), ctypes.byref(code),
                           ctypes.byref(size), log, len(log))
    if rc != 0:
        raise UfoError(f"hipRTC compile failed:\n{lib().ufo_last_error().decode(errors='replace')}")
    try:
        return ctypes.string_at(code, size.value)
    finally:
        lib().ufo_code_free(code)


class Module:
    """A hipRTC-compiled module loaded on a context's device."""

    def __init__(self, ctx: Context, src: str, options: str = ""):
        h = ctypes.c_void_p()
        log = ctypes.create_string_buffer(1 << 16)
        rc = lib().ufo_module_compile(ctx.handle, src.encode(), options.encode(), None, ctypes.byref(h),
                                      log, len(log))
        if rc != 0:
            raise UfoError(f"hipRTC build failed:\n{lib().ufo_last_error().decode(errors='replace')}")
        self.handle = h
        self.ctx = ctx
        self.src = src

    def info(self, kernel: str):
        v, s, p, t = (ctypes.c_int() for _ in range(4))
        check(lib().ufo_kernel_info(self.handle, kernel.encode(), ctypes.byref(v), ctypes.byref(s),
                                    ctypes.byref(p), ctypes.byref(t)), "ufo_kernel_info")
        return {"vgprs": v.value, "private_bytes": p.value, "max_threads": t.value}

    @staticmethod
    def _args(args):
        """Kernel argument array: device tensors by pointer (they must be
        contiguous and on the GPU — the kernels index them densely), Python
        ints as 32-bit ``int`` (range-checked), floats as ``double``."""
        holders = []
        for a in args:
            if isinstance(a, torch.Tensor):
                if a.device.type != "cuda":
                    raise ValueError(f"kernel argument on {a.device}: device tensors only")
                if not a.is_contiguous():
                    raise ValueError("kernel argument tensor is not contiguous")
                holders.append(ctypes.c_void_p(a.data_ptr()))
            elif isinstance(a, bool):
                holders.append(ctypes.c_int(int(a)))
            elif isinstance(a, int):
                if not -2 ** 31 <= a < 2 ** 31:
                    raise ValueError(f"integer kernel argument {a} does not fit a 32-bit int")
                holders.append(ctypes.c_int(a))
            elif isinstance(a, float):
                holders.append(ctypes.c_double(a))
            elif isinstance(a, (ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_float,
                                ctypes.c_uint, ctypes.c_int64, ctypes.c_uint64)):
                holders.append(a)
            else:
                raise TypeError(f"unsupported kernel argument {type(a)}")
        arr = (ctypes.c_void_p * max(1, len(holders)))()
        for i, h in enumerate(holders):
            arr[i] = ctypes.cast(ctypes.byref(h), ctypes.c_void_p)
        return arr, holders

    @staticmethod
    def _cfg(grid, block, shared=0):
        g = tuple(grid) + (1,) * (3 - len(tuple(grid))) if isinstance(grid, (tuple, list)) else (grid, 1, 1)
        b = tuple(block) + (1,) * (3 - len(tuple(block))) if isinstance(block, (tuple, list)) else (block, 1, 1)
        cfg = _Cfg()
        for i in range(3):
            cfg.grid[i] = int(g[i])
            cfg.block[i] = int(b[i])
        cfg.shared_bytes = int(shared)
        return cfg

    def launch(self, kernel: str, grid, block, args, stream=None, shared=0):
        """Asynchronous launch on ``stream`` (default: torch's current
        stream of the context's device)."""
        if stream is None:
            stream = torch.cuda.current_stream(self.ctx.torch_device).cuda_stream
        arr, holders = self._args(args)
        cfg = self._cfg(grid, block, shared)
        check(lib().ufo_launch(self.handle, kernel.encode(), ctypes.byref(cfg), arr, len(holders),
                               ctypes.c_void_p(stream)), f"launch {kernel}")

    def launch_noalloc(self, kernel: str, grid, block, args, scratch_arg: int, scratch_bytes: int,
                       stream=None, shared=0):
        """As ``launch``, with argument ``scratch_arg`` (a placeholder,
        e.g. None) replaced by the library's own device scratch array, of
        which the first ``scratch_bytes`` are zeroed before the launch; no
        device memory is allocated.  Calls that overlap on one device share
        that array and race."""
        if stream is None:
            stream = torch.cuda.current_stream(self.ctx.torch_device).cuda_stream
        args = [ctypes.c_void_p(0) if i == scratch_arg or a is None else a for i, a in enumerate(args)]
        arr, holders = self._args(args)
        cfg = self._cfg(grid, block, shared)
        check(lib().ufo_launch_noalloc(self.handle, kernel.encode(), ctypes.byref(cfg), arr, len(holders),
                                       scratch_arg, scratch_bytes, ctypes.c_void_p(stream)),
              f"launch {kernel} (no allocation)")

    def time(self, kernel: str, grid, block, args, reps=1, stream=None, shared=0) -> float:
        """Milliseconds of ``reps`` back-to-back launches, HIP events on the
        launch stream."""
        if stream is None:
            stream = torch.cuda.current_stream(self.ctx.torch_device).cuda_stream
        arr, holders = self._args(args)
        cfg = self._cfg(grid, block, shared)
        ms = ctypes.c_float()
        check(lib().ufo_time_launch(self.handle, kernel.encode(), ctypes.byref(cfg), arr, len(holders),
                                    ctypes.c_void_p(stream), reps, ctypes.byref(ms)), f"time {kernel}")
        return ms.value

    def __del__(self):
        try:
            if _lib is not None and self.handle:
                _lib.ufo_module_destroy(self.handle)
        except Exception:
            pass


class Graph:
    """A captured sequence of launches on one stream, replayable."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream):
        check(lib().ufo_graph_launch(self.handle, ctypes.c_void_p(stream)), "graph launch")

    def __del__(self):
        try:
            if _lib is not None and self.handle:
                _lib.ufo_graph_destroy(self.handle)
        except Exception:
            pass


def capture(stream, fn_) -> Graph:
    """Run ``fn_()`` under stream capture and return the instantiated graph
    (capture always ends; a graph from a failed ``fn_`` is discarded)."""
    check(lib().ufo_graph_begin(ctypes.c_void_p(stream)), "graph begin")
    h = ctypes.c_void_p()
    try:
        fn_()
    except BaseException:
        if lib().ufo_graph_end(ctypes.c_void_p(stream), ctypes.byref(h)) == 0:
            lib().ufo_graph_destroy(h)
        raise
    check(lib().ufo_graph_end(ctypes.c_void_p(stream), ctypes.byref(h)), "graph end")
    return Graph(h)
