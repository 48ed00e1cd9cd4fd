"""Build libtpamr_hip.so (HIP kernels for gfx950 + host hierarchy builder + C ABI).

Run as ``python twophaseamr.jl_amd/build.py``; the library lands in
``twophaseamr.jl_amd/lib/``. Object files are cached under ``_build/`` and
rebuilt when a source or header is newer.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libtpamr_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
          f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
# packed FP32 instructions stay disabled in device code
DEVICE_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-munsafe-fp-atomics"]
SOURCES = ["hier.cpp", "kernels.hip", "staggered.hip", "momentum.cpp", "cmdgfm.hip", "cmdgfm.cpp", "runtime.cpp"]
HEADERS = ["hier.hpp", "kernels.hpp", "tile.hpp", "staggered.hpp", "momentum.hpp", "cmdgfm.hpp"]


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def build(force: bool = False, verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    headers = [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(ROOT, "include", "tpamr.h")]
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OBJDIR, src + ".o")
        objs.append(obj)
        if force or _newer(obj, [path] + headers):
            if src.endswith(".hip"):
                cmd = [HIPCC, "-x", "hip", *DEVICE_FLAGS, *COMMON, "-c", path, "-o", obj]
            else:
                # host-only translation units (no device code)
                cmd = [HIPCC, *COMMON, "-D__HIP_PLATFORM_AMD__", "-c", path, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
    if force or _newer(LIB, objs):
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB, *objs, "-L/opt/rocm/lib", "-lrocblas"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
    print(LIB)
