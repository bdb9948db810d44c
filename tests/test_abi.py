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


def test_product_fails_loudly_without_the_library(monkeypatch, tmp_path):
    """No CPU fallback: with libufo_hip.so absent, loading the device layer
    raises (and so does every simulation's run(), which goes through it)."""
    from ufo_amd import runtime
    monkeypatch.setattr(runtime, "_lib", None)
    monkeypatch.setattr(runtime, "LIB", str(tmp_path / "libufo_hip.so"))
    with pytest.raises(runtime.UfoError, match="missing"):
        runtime.lib()
    with pytest.raises(runtime.UfoError):
        runtime.device_count()


def test_run_without_a_device_raises():
    """On a machine with no HIP device the drivers raise instead of
    computing anything on the CPU."""
    import torch
    import ufo_amd as ufo
    from ufo_amd import lattices, runtime
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    tr = ufo.Track(lattices.fodo_ring(2), flags=ufo.KICK | ufo.DOUBLE_PRECISION, turns=2, particles=4,
                   parameters=["x"], where=[-1])
    with pytest.raises(runtime.UfoError, match="no HIP device"):
        tr.run()
    sa = ufo.StableAperture(lattices.fodo_ring(2), flags=ufo.KICK | ufo.DOUBLE_PRECISION, turns=2, particles=4,
                            parameters=["x"])
    with pytest.raises(runtime.UfoError, match="no HIP device"):
        sa.run()


@pytest.mark.parametrize("damage", ["truncated", "garbage", "empty"])
def test_damaged_cache_entry_is_rebuilt(monkeypatch, tmp_path, damage):
    """A code-object cache entry that is cut short, overwritten or empty is
    not handed to the loader: the source is compiled again and the entry
    replaced (CPU: hipRTC needs no device)."""
    from ufo_amd import runtime
    monkeypatch.setenv("UFO_CACHE_DIR", str(tmp_path))
    src = 'extern "C" __global__ void k(double *a) { a[threadIdx.x] *= 2.0; } // cache entry test'
    good = runtime.compile_only(src)
    (entry,) = tmp_path.glob("*.hsaco")
    entry.write_bytes({"truncated": good[: len(good) // 2], "garbage": b"\x7fELF" + bytes(200),
                       "empty": b""}[damage])
    assert runtime.compile_only(src) == good
    assert entry.read_bytes() == good


def test_noalloc_launch_argument_errors():
    """ufo_launch_noalloc refuses a scratch size outside 1..UFO_SCRATCH_MAX_BYTES
    and a scratch slot outside the argument list before it touches a device."""
    from ufo_amd import runtime
    lib = runtime.lib()
    assert lib.ufo_launch_noalloc(None, b"k", None, None, 0, 0, (1 << 20) + 1, None) == -1
    assert b"scratch of 1048577 bytes" in lib.ufo_last_error()
    assert lib.ufo_launch_noalloc(None, b"k", None, None, 0, 0, 0, None) == -1
    assert b"scratch of 0 bytes" in lib.ufo_last_error()
    assert lib.ufo_launch_noalloc(None, b"k", None, None, 0, 0, 4, None) == -1
    assert b"null argument" in lib.ufo_last_error()
    header = open(os.path.join(ROOT, "include", "ufo_hip.h")).read()
    assert "#define UFO_SCRATCH_MAX_BYTES (1u << 20)" in header
