"""Build of libufo_hip.so (the C-ABI device layer) in-tree with hipcc.

The generated tracking kernels are not built here: they are compiled at run
time by hipRTC through the library (ufo_compile / ufo_module_compile).
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libufo_hip.so")
SRC = os.path.join(HERE, "csrc", "ufo_runtime.cpp")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# No packed-FP32 instructions in any HIP build (they have misbehaved on the
# target machines); the flag is also passed to every hipRTC kernel build.
NO_PACKED_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build_library(verbose=False) -> str:
    tmp = LIB + f".tmp{os.getpid()}"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared", *NO_PACKED_FP32,
           "-Wall", "-o", tmp, SRC, "-lhiprtc", "-lamdhip64"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
