"""The C ABI: the library loads without a GPU and exports every entry point
declared in include/tpamr.h; a host-only context reports device calls as
errors instead of computing them elsewhere."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import tpamr
from tpamr import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tpamr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tpamr_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name


def test_host_only_context_refuses_compute():
    m = tpamr.BlockTreeMesh(block_size=(8, 8))
    ctx = _lib.Context(-1)
    m.attach(ctx)
    k = np.ones(m.nfaces)
    _lib.check(_lib.lib().tpamr_set_face_coefficients(ctx.h, _lib.ptr(k)), ctx.h)
    x = np.zeros(m.ncells)
    st = _lib.Stats()
    rc = _lib.lib().tpamr_solve(ctx.h, _lib.ptr(x), _lib.ptr(x), 1e-10, 0.0, 10, 0, C.byref(st))
    assert rc == -5 and b"host-only" in _lib.lib().tpamr_last_error(ctx.h)
    rc = _lib.lib().tpamr_apply(ctx.h, _lib.ptr(x), _lib.ptr(x))
    assert rc == -5


def test_host_only_context_refuses_momentum():
    m = tpamr.BlockTreeMesh(base_mesh=(16, 16), block_size=(8, 8))
    ctx = _lib.Context(-1)
    m.attach(ctx)
    from tpamr import navier_stokes as ns
    rc = ns._lib().tpamr_set_momentum(ctx.h, 0.01, 2)
    assert rc == -5 and b"host-only" in _lib.lib().tpamr_last_error(ctx.h)
    u = np.zeros(m.nfaces)
    rc = ns._lib().tpamr_advect(ctx.h, _lib.ptr(u), _lib.ptr(u), 0.1)
    assert rc == -5
