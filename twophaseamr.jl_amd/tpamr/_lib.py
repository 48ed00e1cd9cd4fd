"""ctypes binding of libtpamr_hip.so (the C ABI declared in include/tpamr.h).

The library must exist: there is no CPU fallback. Build it with
``python twophaseamr.jl_amd/build.py`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libtpamr_hip.so")

TPAMR_FP64 = 0
TPAMR_FP32_VCYCLE = 1
TPAMR_FP32_REFINE = 2

ERRORS = {
    -1: ValueError,       # ArgumentError
    -2: ValueError,       # DimensionMismatch
    -3: ValueError,       # non-finite input
    -4: RuntimeError,     # device / HIP runtime
    -5: RuntimeError,     # call order
    -6: ValueError,       # incompatible rhs with compatibility = :error
}


class TpamrError(RuntimeError):
    pass


class DimensionMismatch(ValueError):
    pass


class Stats(C.Structure):
    _fields_ = [("cycles", C.c_int), ("converged", C.c_int), ("compatible", C.c_int), ("reason", C.c_int),
                ("initial_l2", C.c_double), ("initial_linf", C.c_double), ("final_l2", C.c_double),
                ("final_linf", C.c_double), ("target_linf", C.c_double), ("residual_scale", C.c_double),
                ("removed_rhs_mean", C.c_double), ("solve_ms", C.c_double), ("krylov_runs", C.c_int),
                ("replacements", C.c_int)]


class MGOptions(C.Structure):
    _fields_ = [("coarse_size", C.c_int), ("pre_sweeps", C.c_int), ("post_sweeps", C.c_int),
                ("max_levels", C.c_int)]


_lib = None

_SIGNATURES = {
    "tpamr_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int]),
    "tpamr_destroy": (None, [C.c_void_p]),
    "tpamr_last_error": (C.c_char_p, [C.c_void_p]),
    "tpamr_set_mesh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_int, C.c_int, C.c_void_p]),
    "tpamr_mesh_sizes": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tpamr_mesh_blocks": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tpamr_mesh_faces": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_mesh_boundary_faces": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_set_multigrid": (C.c_int, [C.c_void_p, C.POINTER(MGOptions)]),
    "tpamr_set_face_coefficients": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tpamr_num_levels": (C.c_int, [C.c_void_p]),
    "tpamr_level_info": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "tpamr_level_matrix": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "tpamr_level_transfer": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "tpamr_level_kappa": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "tpamr_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int,
                              C.POINTER(Stats)]),
    "tpamr_solve_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int,
                                     C.c_int, C.POINTER(Stats)]),
    "tpamr_solve_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int,
                                             C.c_int, C.POINTER(Stats)]),
    "tpamr_vcycle": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_krylov": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p,
                               C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "tpamr_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_apply_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_mixture_face_coefficients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p]),
    "tpamr_surface_tension_jumps": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                              C.c_void_p]),
    "tpamr_surface_tension_jumps_cells": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                                    C.c_int, C.c_void_p, C.c_void_p]),
    "tpamr_weighted_jump_force": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_gravity_jumps": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_void_p]),
    "tpamr_gravity": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_void_p]),
    "tpamr_divergence": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_weighted_gradient": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_halo_sizes": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "tpamr_halo_schedule": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "tpamr_fill_halo": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_project": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int,
                                C.c_void_p, C.POINTER(Stats)]),
    "tpamr_pressure_predictor": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
    "tpamr_project_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int,
                                   C.c_void_p, C.c_double, C.c_void_p, C.POINTER(Stats)]),
    "tpamr_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p,
                                C.c_void_p, C.POINTER(Stats)]),
    "tpamr_project_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                        C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Stats)]),
    "tpamr_set_face_coefficients_device": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tpamr_set_caller_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tpamr_hierarchy_reuse": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "tpamr_time_stream": (C.c_int, [C.c_void_p, C.c_longlong, C.c_int, C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
    "tpamr_apply_force_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    "tpamr_divergence_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_mixture_face_coefficients_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                                         C.c_void_p]),
    "tpamr_gravity_jumps_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                             C.c_double, C.c_double, C.c_double, C.c_void_p]),
    "tpamr_project_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int,
                                       C.c_void_p, C.c_double, C.c_void_p, C.POINTER(Stats)]),
    "tpamr_project_adjoint_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                               C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.POINTER(Stats)]),
    "tpamr_time_smoother": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tpamr_set_pressure_boundary": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_weighted_gradient_boundary": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_project_open": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                     C.c_int, C.c_void_p, C.c_void_p]),
    "tpamr_level_boundary": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "tpamr_component_means": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_void_p, C.c_void_p]),
    "tpamr_mixture_boundary_coefficients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                                      C.c_void_p]),
    "tpamr_level_components": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p, C.c_void_p]),
}


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with python twophaseamr.jl_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f64(a, n=None, name="array"):
    arr = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and arr.shape != (n,):
        raise DimensionMismatch(f"{name} must have length {n}, got shape {arr.shape}")
    return arr


def order_after_torch(handle, device):
    """Make the library's device-memory entry points on this context wait for
    the work queued so far on torch's current stream of `device` (where the
    caller wrote their inputs)."""
    import torch
    stream = torch.cuda.current_stream(device)
    check(lib().tpamr_set_caller_stream(handle, C.c_void_p(stream.cuda_stream)), handle)


def check(rc, handle):
    if rc != 0:
        msg = lib().tpamr_last_error(handle)
        text = msg.decode() if msg else "unknown error"
        if rc == -2:
            raise DimensionMismatch(text)
        raise ERRORS.get(rc, TpamrError)(text)


class Context:
    """Owns one tpamr_ctx (device >= 0, or -1 for a host-only context)."""

    def __init__(self, device=0, precision=TPAMR_FP64):
        h = C.c_void_p()
        rc = lib().tpamr_create(C.byref(h), int(device), int(precision))
        self.h = h
        self.device = device
        self.precision = precision
        if rc != 0:
            try:
                check(rc, h)
            finally:
                lib().tpamr_destroy(h)
                self.h = None

    def close(self):
        if getattr(self, "h", None):
            lib().tpamr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
