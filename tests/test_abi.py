"""The C-ABI library loads and exports every entry point include/ufo_hip.h
declares; error paths that need no device.  CPU only."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def header_functions():
    text = open(os.path.join(ROOT, "include", "ufo_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(ufo_\w+)\s*\(", text, re.M)))


def test_header_lists_entry_points():
    names = header_functions()
    assert "ufo_module_compile" in names and "ufo_launch" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    from ufo_amd import runtime
    lib = runtime.lib()
    for name in header_functions():
        assert hasattr(lib, name), name


def test_no_device_errors_and_compile_without_device():
    import torch
    from ufo_amd import runtime
    lib = runtime.lib()
    n = runtime.device_count()
    if not torch.cuda.is_available():
        assert n == 0
        h = ctypes.c_void_p()
        assert lib.ufo_context_create(0, ctypes.byref(h)) != 0
        assert b"no device" in lib.ufo_last_error()
    assert lib.ufo_launch(None, b"k", None, None, 0, None) == -1
    code = runtime.compile_only('extern "C" __global__ void k(double *a) { a[threadIdx.x] += 1.0; }')
    assert code[:4] in (b"\x7fELF", b"__CL")
