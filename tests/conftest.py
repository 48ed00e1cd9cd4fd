"""Shared test setup: markers, import paths and mesh fixtures."""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the oracle's OpenMP loops: idle threads yield instead of spinning, and
# parallel test workers share the host's cores
os.environ.setdefault("OMP_WAIT_POLICY", "passive")
if os.environ.get("PYTEST_XDIST_WORKER"):
    os.environ.setdefault("OMP_NUM_THREADS", "2")
sys.path.insert(0, os.path.join(ROOT, "twophaseamr.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: larger CPU cases")
    # torch carries its own HIP runtime; it finds the GPU only when it
    # initialises before the library's runtime has opened the device, so GPU
    # sessions (which hand torch tensors to the library) start it first
    markexpr = getattr(config.option, "markexpr", "") or ""
    if "gpu" in markexpr and "not gpu" not in markexpr:
        gpu_available()


def gpu_available():
    try:
        import torch  # noqa: F401
        return torch.cuda.is_available()
    except Exception:
        return False
