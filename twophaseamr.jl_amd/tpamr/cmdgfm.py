"""CMDGFM reduced-pressure/jump block solve over the C ABI: the linear solve of
the reference's two-velocity CMDGFM projection on its assembled system.

Reference interface (TwoPhaseAMR.jl, src/two_phase/projection/):
- the assembled system of ``_assemble_cmdgfm_system!``
  (cmdgfm_projection.jl:1484-1523): ``matrix`` over the reduced pressures
  (``pressure_indices`` minus ``pressure_anchors``) and the mixed-face jumps,
  ``right_hand_side``, ``jump_pressure`` (its C block) and the shared
  pressure hierarchy (``shared_multigrid``);
- ``CMDGFMBlockPreconditioner`` with the reduced-pressure block
  (``_cmdgfm_pressure_preconditioner``, :1632-1700, ``ldiv!`` :374-440) and
  the ILU of the jump Schur complement (``_cmdgfm_jump_schur!``,
  cmdgfm_cache.jl:550-610; ``jump_drop_tolerance`` 1e-3 by default);
- the solve of ``_project_cmdgfm_assembled!`` (cmdgfm_projection.jl:
  2201-2318) and the solve fields of ``CMDGFMProjectionStats``.

The pressure block's V-cycle is the EllipticSolver's hierarchy: its
operator's face coefficients must be those of the system's pressure
operator. Assembling the system (interface geometry, jump stencils) and the
velocity correction after the solve are the caller's.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from .hierarchy import HostHierarchy
from .solver import EllipticSolver


class CmdgfmStats(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int), ("krylov_runs", C.c_int),
                ("initial_guess_used", C.c_int), ("pressure_unknowns", C.c_int), ("jump_unknowns", C.c_int),
                ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("pressure_defect", C.c_double), ("jump_defect", C.c_double), ("anchor_defect", C.c_double),
                ("asymmetry", C.c_double), ("jump_diagonal_margin", C.c_double)]


_P = C.c_void_p
_SIGS = {
    "tpamr_cmdgfm_set_system": (C.c_int, [_P, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P, _P, C.c_int, _P, _P, _P, _P,
                                          _P, C.c_double]),
    "tpamr_cmdgfm_jump_factor": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "tpamr_cmdgfm_solve": (C.c_int, [_P, _P, _P, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _P, C.c_int,
                                     C.POINTER(CmdgfmStats)]),
    "tpamr_cmdgfm_solve_device": (C.c_int, [_P, _P, _P, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _P,
                                            C.c_int, C.POINTER(CmdgfmStats)]),
}


def _lib():
    lib = L.lib()
    if not getattr(lib, "_cmdgfm_bound", False):
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib._cmdgfm_bound = True
    return lib


@dataclass
class CMDGFMSolveStats:
    """The solve fields of CMDGFMProjectionStats."""
    converged: bool
    iterations: int
    krylov_runs: int
    initial_guess_used: bool
    pressure_unknowns: int
    jump_unknowns: int
    initial_residual: float
    final_residual: float
    pressure_defect: float
    jump_defect: float
    anchor_defect: float
    asymmetry: float
    jump_diagonal_margin: float
    preconditioner: str

    @classmethod
    def _from(cls, st: CmdgfmStats, preconditioner):
        return cls(bool(st.converged), st.iterations, st.krylov_runs, bool(st.initial_guess_used),
                   st.pressure_unknowns, st.jump_unknowns, st.initial_residual, st.final_residual,
                   st.pressure_defect, st.jump_defect, st.anchor_defect, st.asymmetry, st.jump_diagonal_margin,
                   preconditioner)


def _csr(matrix, shape=None, name="matrix"):
    m = sp.csr_matrix(matrix, dtype=np.float64)
    if shape is not None and m.shape != shape:
        raise L.DimensionMismatch(f"{name} must have shape {shape}, got {m.shape}")
    m.sum_duplicates()
    m.sort_indices()
    return (np.ascontiguousarray(m.indptr, np.int32), np.ascontiguousarray(m.indices, np.int32),
            np.ascontiguousarray(m.data, np.float64))


def _sync(owner):
    fn = getattr(owner, "_sync", None)
    if fn is not None:
        fn()


def _ncells(owner):
    return owner.operator.mesh.ncells if hasattr(owner, "operator") else owner.mesh.ncells


def _i32(a, n, name):
    arr = np.ascontiguousarray(a, dtype=np.int32)
    if arr.shape != (n,):
        raise L.DimensionMismatch(f"{name} must have length {n}")
    return arr


class CMDGFMBlockSystem:
    """The assembled reduced-pressure/jump system of one CMDGFM projection on
    ``solver``'s context (the solver's hierarchy is the shared pressure
    multigrid; a HostHierarchy gives a host-only system whose jump factor can
    be inspected).

    ``matrix``: (np + nx) square, unknowns 0..np-1 the reduced pressures of
    cells ``pressure_indices`` (0-based) relative to ``pressure_anchors``,
    the rest the mixed-face jumps. ``pressure_volumes``: the retained cells'
    volumes; ``jump_lengths``: each mixed face's defect length. ``omitted``
    (rows over the same unknowns), ``omitted_rhs`` and ``omitted_volumes``:
    the anchor and isolated cells' equations for the anchor defect (None:
    not checked)."""

    def __init__(self, solver: EllipticSolver | HostHierarchy, matrix, pressure_indices, pressure_anchors,
                 pressure_volumes, jump_lengths, *, omitted=None, omitted_rhs=None, omitted_volumes=None,
                 drop_tolerance=1e-3):
        n = int(matrix.shape[0])
        np_ = len(pressure_indices)
        nx = n - np_
        if nx < 0:
            raise L.DimensionMismatch("more pressure indices than unknowns")
        self.solver = solver
        self.ctx = solver.ctx
        self.np, self.nx, self.n = np_, nx, n
        self._a = _csr(matrix, (n, n))
        self._idx = _i32(pressure_indices, np_, "pressure_indices")
        self._anc = _i32(pressure_anchors, np_, "pressure_anchors")
        self._vol = L.f64(pressure_volumes, np_, "pressure_volumes")
        self._len = L.f64(jump_lengths, nx, "jump_lengths")
        if omitted is not None:
            m = int(omitted.shape[0])
            self._om = _csr(omitted, (m, n), "omitted")
            self._orhs = L.f64(omitted_rhs, m, "omitted_rhs")
            self._ovol = L.f64(omitted_volumes, m, "omitted_volumes")
        else:
            m = 0
            self._om = (np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
            self._orhs = self._ovol = np.zeros(0)
        _sync(solver)
        a, om = self._a, self._om
        L.check(_lib().tpamr_cmdgfm_set_system(
            self.ctx.h, np_, nx, L.ptr(a[0]), L.ptr(a[1]), L.ptr(a[2]), L.ptr(self._idx), L.ptr(self._anc),
            L.ptr(self._vol), L.ptr(self._len), m, L.ptr(om[0]), L.ptr(om[1]), L.ptr(om[2]), L.ptr(self._orhs),
            L.ptr(self._ovol), float(drop_tolerance)), self.ctx.h)

    def jump_factor(self):
        """(S, L, U): the jump Schur complement J - C D^-1 B, the strictly
        lower unit factor and the upper factor (with its diagonal) of its
        Crout ILU, as CSR matrices."""
        sizes = np.zeros(4, np.int32)
        lib = _lib()
        L.check(lib.tpamr_cmdgfm_jump_factor(self.ctx.h, L.ptr(sizes), *([None] * 10)), self.ctx.h)
        nx, ns, nl, nu = (int(v) for v in sizes)
        if nx == 0:
            z = sp.csr_matrix((0, 0))
            return z, z, z
        sptr, lptr, uptr = (np.zeros(nx + 1, np.int32) for _ in range(3))
        scol, lcol, ucol = np.zeros(ns, np.int32), np.zeros(nl, np.int32), np.zeros(nu, np.int32)
        sval, lval, uval, udiag = np.zeros(ns), np.zeros(nl), np.zeros(nu), np.zeros(nx)
        L.check(lib.tpamr_cmdgfm_jump_factor(self.ctx.h, L.ptr(sizes), L.ptr(sptr), L.ptr(scol), L.ptr(sval),
                                             L.ptr(lptr), L.ptr(lcol), L.ptr(lval), L.ptr(uptr), L.ptr(ucol),
                                             L.ptr(uval), L.ptr(udiag)), self.ctx.h)
        s = sp.csr_matrix((sval, scol, sptr), shape=(nx, nx))
        lo = sp.csr_matrix((lval, lcol, lptr), shape=(nx, nx))
        up = sp.csr_matrix((uval, ucol, uptr), shape=(nx, nx)) + sp.diags(udiag)
        return s, lo, sp.csr_matrix(up)

    def solve(self, rhs, *, dt, density, tolerance=float(np.sqrt(np.finfo(float).eps)), max_iterations=500,
              warm_start=True, previous_increment=None, preconditioner="block"):
        """The solve of the CMDGFM projection on this system: returns the
        increment (reduced pressures, then jumps) and its stats.
        ``preconditioner``: 'block' (the reference's default) or None."""
        pc = self._pc(preconditioner)
        b = L.f64(rhs, self.n, "right-hand side")
        prev = None if previous_increment is None else L.f64(previous_increment, _ncells(self.solver),
                                                              "previous_increment")
        x = np.zeros(self.n)
        st = CmdgfmStats()
        _sync(self.solver)
        L.check(_lib().tpamr_cmdgfm_solve(self.ctx.h, L.ptr(b), L.ptr(x), float(dt), float(density),
                                          float(tolerance), int(max_iterations), int(bool(warm_start)), L.ptr(prev),
                                          pc, C.byref(st)), self.ctx.h)
        return x, CMDGFMSolveStats._from(st, "cmdgfm_block" if pc else "none")

    def solve_device(self, rhs, solution, *, dt, density, tolerance=float(np.sqrt(np.finfo(float).eps)),
                     max_iterations=500, warm_start=True, previous_increment=None, preconditioner="block"):
        """solve with torch float64 device tensors; `solution` is overwritten."""
        import torch

        pc = self._pc(preconditioner)
        for t, n, what in ((rhs, self.n, "rhs"), (solution, self.n, "solution")):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.shape == (n,)
                    and t.is_contiguous()):
                raise L.DimensionMismatch(f"{what} must be a contiguous float64 device tensor of length {n}")
        prev = previous_increment
        if prev is not None and not (isinstance(prev, torch.Tensor) and prev.is_cuda and prev.dtype == torch.float64
                                     and prev.shape == (_ncells(self.solver),) and prev.is_contiguous()):
            raise L.DimensionMismatch("previous_increment must be a contiguous float64 device cell tensor")
        st = CmdgfmStats()
        _sync(self.solver)
        L.order_after_torch(self.ctx.h, self.ctx.device)
        L.check(_lib().tpamr_cmdgfm_solve_device(
            self.ctx.h, C.c_void_p(rhs.data_ptr()), C.c_void_p(solution.data_ptr()), float(dt), float(density),
            float(tolerance), int(max_iterations), int(bool(warm_start)),
            None if prev is None else C.c_void_p(prev.data_ptr()), pc, C.byref(st)), self.ctx.h)
        return CMDGFMSolveStats._from(st, "cmdgfm_block" if pc else "none")

    @staticmethod
    def _pc(preconditioner):
        if preconditioner not in ("block", None):
            raise ValueError("preconditioner must be 'block' or None")
        return 1 if preconditioner == "block" else 0
