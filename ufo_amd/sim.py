"""Shared plumbing of the simulation drivers: precision, host/device buffers,
kernel build through the context, launch sizing."""
from __future__ import annotations

import numpy as np
import torch

from . import codegen
from .flags import DOUBLE_PRECISION


def hip_options(options) -> str:
    """hipRTC options of a kernel build.  OpenCL options of the reference
    (``-cl-...``, ufo/__init__.py:22) have no meaning here and are dropped:
    the kernels keep IEEE semantics and use explicit fused multiply-adds."""
    import ufo_amd
    extra = options if options is not None else ufo_amd.DEFAULT_CL_OPTIONS
    kept = [o for o in (extra or "").split() if not o.startswith("-cl-")]
    return " ".join([codegen.HIPRTC_OPTIONS] + kept)


def has_gpu() -> bool:
    return torch.cuda.is_available()


def host_buffer(shape, dtype):
    """Host tensor for results/inputs: pinned when a GPU is present."""
    return torch.zeros(shape, dtype=dtype, pin_memory=has_gpu())


class Simulation:
    """Base of Track/StableAperture/Optics: owns the generated kernel, its
    module on the context's device and the launch sizing."""

    def _init_common(self, flags, context, options):
        self.flags = flags
        self.fp32 = not (flags & DOUBLE_PRECISION)
        self.np_dtype = np.float32 if self.fp32 else np.float64
        self.dtype = torch.float32 if self.fp32 else torch.float64
        self._context = context
        self.options = hip_options(options)
        self._module = None

    @property
    def ctx(self):
        if self._context is None:
            from . import runtime
            self._context = runtime.default_context()
        return self._context

    @property
    def device(self):
        return self.ctx.torch_device

    def module(self):
        if self._module is None:
            self._module = self.ctx.module(self.kernel.src, self.options)
        return self._module

    def compile_only(self, arch="gfx950") -> bytes:
        """Build the kernel with hipRTC for ``arch`` without a device."""
        from . import runtime
        return runtime.compile_only(self.kernel.src, self.options, arch)

    @property
    def src(self):
        return self.kernel.src

    def resident_lanes(self, block=None, waves_per_simd=8) -> int:
        """Lanes the device holds at once for this kernel with at most
        ``waves_per_simd`` waves per SIMD (occupancy from the kernel's VGPR
        count, 8 waves per SIMD at most, 4 SIMDs per CU)."""
        info = self.module().info(self.kernel.name)
        v = max(8, info["vgprs"])
        v = (v + 7) // 8 * 8
        waves = max(1, min(8, waves_per_simd, 512 // v))
        return self.ctx.device.compute_units * 4 * waves * 64

    def stream(self):
        return torch.cuda.current_stream(self.device)

    def _host_pool(self):
        """The pinned pool tensor holding the current ``parameters``.  As in
        the reference, whatever array ``parameters`` refers to at ``run()``
        is what is tracked: filled in place (the usual form) it already is
        the pinned buffer; if the attribute was rebound to another array, that
        array is copied in (it must have the pool's shape)."""
        host = self._pool_host
        view = host.numpy()
        p = self.parameters
        same = (isinstance(p, np.ndarray) and p.dtype == view.dtype and p.shape == view.shape
                and p.ctypes.data == view.ctypes.data)
        if not same:
            a = np.asarray(p, dtype=view.dtype)
            if a.shape != view.shape:
                if a.size == view.size:
                    a = a.reshape(view.shape)
                else:
                    raise ValueError(f"parameters has shape {a.shape}; the pool is {view.shape}")
            view[...] = a
        return host
