"""Pressure solver surface mirroring TwoPhaseAMR.jl's linear-solver API.

Reference interfaces replaced here (file:line in TwoPhaseAMR.jl):
  MultigridOptions       src/solvers/multigrid/options.jl:1-52
  LinearSolverOptions    src/solvers/linear_options.jl:10-47
  PoissonOperator        src/operators/elliptic.jl:969-980 (face_coefficient form;
                         velocity_boundary / boundary_face_coefficient, :683-800)
  VelocityBoundaryConditions, NoSlip, FreeSlip, Outflow, ZeroGradientVelocity
                         (the open pressure sides they imply, elliptic.jl:76-88)
  EllipticSolver         src/solvers/multigrid/elliptic_solver.jl:282-302
  solve!                 src/solvers/multigrid/elliptic_solver.jl:628-675
  ldiv!(z, P, r)         src/solvers/multigrid/cycles.jl:275-283
  LinearSolverStats      src/solvers/multigrid/stats.jl:14-36
All computation runs in libtpamr_hip.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .mesh import BlockTreeMesh

REASONS = {0: "none", 1: "solved", 2: "strong residual target", 3: "zero curvature", 4: "iteration limit",
           5: "indefinite preconditioner"}


class MultigridOptions:
    """MultigridOptions(; coarse_size=64, pre_sweeps=2, post_sweeps=2, max_levels=32)."""

    def __init__(self, coarse_size=64, pre_sweeps=2, post_sweeps=2, max_levels=32):
        vals = (coarse_size, pre_sweeps, post_sweeps, max_levels)
        if not all(isinstance(v, (int, np.integer)) and not isinstance(v, bool) for v in vals):
            raise ValueError("multigrid sizes and sweep counts must be integers")
        if coarse_size <= 0:
            raise ValueError("coarse_size must be positive")
        if pre_sweeps < 0:
            raise ValueError("pre_sweeps must be nonnegative")
        if post_sweeps < 0:
            raise ValueError("post_sweeps must be nonnegative")
        if max_levels <= 0:
            raise ValueError("max_levels must be positive")
        self.coarse_size = int(coarse_size)
        self.pre_sweeps = int(pre_sweeps)
        self.post_sweeps = int(post_sweeps)
        self.max_levels = int(max_levels)

    def _c(self):
        return L.MGOptions(self.coarse_size, self.pre_sweeps, self.post_sweeps, self.max_levels)

    def __repr__(self):
        return (f"MultigridOptions(coarse_size={self.coarse_size}, pre_sweeps={self.pre_sweeps}, "
                f"post_sweeps={self.post_sweeps}, max_levels={self.max_levels})")


_PRECONDITIONERS = (":auto", ":geometric", ":tree_geometric", ":amg", ":ilu", ":ichol")


class LinearSolverOptions:
    """LinearSolverOptions(; krylov=None, krylov_options={}, preconditioner=':auto',
    preconditioner_options={}). This library implements CG with the tree
    geometric multigrid; `preconditioner_options` may carry `precision`
    ('fp64' or 'fp32_vcycle') and `device`."""

    def __init__(self, krylov=None, krylov_options=None, preconditioner=":auto", preconditioner_options=None):
        if isinstance(krylov, str):
            raise ValueError("krylov must be None, a Krylov workspace, or a workspace constructor; "
                             "symbols are not supported")
        if isinstance(preconditioner, str) and preconditioner not in _PRECONDITIONERS:
            raise ValueError("preconditioner must be :auto, :geometric, :tree_geometric, :amg, :ilu, :ichol, "
                             "None, a factory, or a pre-built preconditioner")
        krylov_options = {} if krylov_options is None else krylov_options
        preconditioner_options = {} if preconditioner_options is None else preconditioner_options
        if not isinstance(krylov_options, dict):
            raise ValueError("krylov_options must be a named tuple")
        if not isinstance(preconditioner_options, dict):
            raise ValueError("preconditioner_options must be a named tuple")
        self.krylov = krylov
        self.krylov_options = dict(krylov_options)
        self.preconditioner = preconditioner
        self.preconditioner_options = dict(preconditioner_options)


@dataclass
class LinearSolverStats:
    cycles: int = 0
    converged: bool = False
    compatible: bool = True
    initial_l2: float = float("inf")
    initial_linf: float = float("inf")
    final_l2: float = float("inf")
    final_linf: float = float("inf")
    target_linf: float = float("inf")
    residual_scale: float = 1.0
    removed_rhs_mean: float = 0.0
    removed_rhs_component_means: list = field(default_factory=list)
    removed_rhs_component_anchors: list = field(default_factory=list)
    method: str = "cg"
    preconditioner: str = "tree_geometric_multigrid"
    reason: str = "none"
    solve_ms: float = 0.0
    krylov_runs: int = 0
    replacements: int = 0

    @classmethod
    def _from(cls, st: L.Stats, precision, handle=None):
        means, anchors = [st.removed_rhs_mean], [1]
        if handle is not None:
            n = C.c_int()
            L.check(L.lib().tpamr_component_means(handle, C.byref(n), None, None), handle)
            roots = np.zeros(max(n.value, 1), np.int32)
            vals = np.zeros(max(n.value, 1))
            L.check(L.lib().tpamr_component_means(handle, C.byref(n), L.ptr(roots), L.ptr(vals)), handle)
            means = [float(v) for v in vals[:n.value]]
            anchors = [int(r) for r in roots[:n.value]]
        return cls(cycles=st.cycles, converged=bool(st.converged), compatible=bool(st.compatible),
                   initial_l2=st.initial_l2, initial_linf=st.initial_linf, final_l2=st.final_l2,
                   final_linf=st.final_linf, target_linf=st.target_linf, residual_scale=st.residual_scale,
                   removed_rhs_mean=st.removed_rhs_mean,
                   removed_rhs_component_means=means, removed_rhs_component_anchors=anchors,
                   preconditioner="tree_geometric_multigrid" + ("" if precision == L.TPAMR_FP64 else "_fp32"),
                   method="cg" if precision != L.TPAMR_FP32_REFINE else "refinement(cg_fp32)",
                   reason=REASONS.get(st.reason, str(st.reason)), solve_ms=st.solve_ms,
                   krylov_runs=st.krylov_runs, replacements=st.replacements)


class NoSlip:
    """Wall: homogeneous Neumann pressure side."""


class FreeSlip:
    """Slip wall: homogeneous Neumann pressure side."""


class Outflow:
    """Open side: homogeneous Dirichlet pressure side."""


class ZeroGradientVelocity:
    """Open side: homogeneous Dirichlet pressure side."""


class VelocityBoundaryConditions:
    """VelocityBoundaryConditions(; left, right, bottom, top): the pressure
    operator reads only which sides are open (Outflow, ZeroGradientVelocity)."""

    def __init__(self, left=None, right=None, bottom=None, top=None):
        self.left = NoSlip() if left is None else left
        self.right = NoSlip() if right is None else right
        self.bottom = NoSlip() if bottom is None else bottom
        self.top = NoSlip() if top is None else top


def pressure_boundary_sides(boundary):
    """`_pressure_boundary_sides` (elliptic.jl:76-84): (left, right, bottom,
    top) open flags of None, a wall condition, a 4-tuple of flags, or
    VelocityBoundaryConditions."""
    if boundary is None or isinstance(boundary, (NoSlip, FreeSlip)):
        return (False, False, False, False)
    if isinstance(boundary, VelocityBoundaryConditions):
        return tuple(isinstance(c, (Outflow, ZeroGradientVelocity))
                     for c in (boundary.left, boundary.right, boundary.bottom, boundary.top))
    sides = tuple(bool(v) for v in boundary)
    if len(sides) != 4:
        raise ValueError("open sides are four flags: left, right, bottom, top")
    return sides


class PoissonOperator:
    """PoissonOperator(mesh; face_coefficient=1, velocity_boundary=None,
    boundary_face_coefficient=None): -div(kappa grad p) in integrated
    finite-volume form with the AMR face Riesz map. Sides are homogeneous
    Neumann (walls) or periodic, except the open sides of `velocity_boundary`
    (homogeneous Dirichlet, boundary distance half the cell spacing).
    `face_coefficient` is a number or a per-face vector in the mesh's face
    order. `boundary_face_coefficient` (one per boundary face) defaults to a
    scalar face coefficient on the open faces, or, for a vector, to the
    area-weighted interior coefficients on the opposite side of each open
    face's cell (computed by the library)."""

    def __init__(self, mesh: BlockTreeMesh, face_coefficient=1.0, velocity_boundary=None,
                 boundary_face_coefficient=None):
        self.mesh = mesh
        self.version = 0
        self.boundary_sides = pressure_boundary_sides(velocity_boundary)
        self.face_coefficient = self._coefficients(face_coefficient)
        if boundary_face_coefficient is None and np.isscalar(face_coefficient) and any(self.boundary_sides):
            open_face = np.asarray([self.boundary_sides[2 * a + (1 if sd > 0 else 0)]
                                    for a, sd in zip(mesh.bface_axis, mesh.bface_side)], bool)
            boundary_face_coefficient = np.where(open_face, float(face_coefficient), 0.0)
        self.boundary_face_coefficient = self._boundary(boundary_face_coefficient)

    def _boundary(self, coefficient):
        if coefficient is None:
            return None
        m = self.mesh
        values = np.asarray(np.broadcast_to(np.asarray(coefficient, np.float64), (m.nboundary_faces,)), np.float64)
        if not np.all(np.isfinite(values)) or np.any(values < 0):
            raise ValueError("boundary coefficients must be finite and nonnegative")
        return np.ascontiguousarray(values)

    def _coefficients(self, coefficient):
        m = self.mesh
        if np.isscalar(coefficient):
            values = np.full(m.nfaces, float(coefficient))
        elif callable(coefficient):
            raise ValueError("pass face coefficients as a vector in face order")
        else:
            values = np.asarray(coefficient, dtype=np.float64)
            if values.shape != (m.nfaces,):
                raise L.DimensionMismatch("one coefficient is required per interior face")
        if not np.all(np.isfinite(values)):
            raise ValueError("face coefficients must be finite")
        if np.any(values < 0):
            raise ValueError("face coefficients must be nonnegative")
        return np.ascontiguousarray(values)

    def replace_face_coefficient(self, face_coefficient, boundary_face_coefficient=None):
        """`_replace_face_coefficient` (elliptic.jl:809-820): new interior
        coefficients and, when given, new boundary coefficients (kept
        otherwise). Solvers built on this operator pick them up at their next
        call; after changing `face_coefficient` in place, call the solver's
        refresh()."""
        self.face_coefficient = self._coefficients(face_coefficient)
        if boundary_face_coefficient is not None:
            self.boundary_face_coefficient = self._boundary(boundary_face_coefficient)
            self.boundary_version = getattr(self, "boundary_version", 0) + 1
        self.version += 1
        return self


def _level_components(ctx, level, ncells):
    n = C.c_int()
    L.check(L.lib().tpamr_level_components(ctx.h, level, C.byref(n), None, None), ctx.h)
    slot = np.zeros(ncells, np.int32)
    roots = np.zeros(max(n.value, 1), np.int32)
    L.check(L.lib().tpamr_level_components(ctx.h, level, C.byref(n), L.ptr(slot), L.ptr(roots)), ctx.h)
    return slot, roots[:n.value]


class EllipticSolver:
    """EllipticSolver(operator; multigrid=MultigridOptions(), precision='fp64',
    device=0): CG preconditioned by one tree-GMG V-cycle per iteration.
    'fp64': FP64 CG, FP64 V-cycle. 'fp32_vcycle': FP64 CG, FP32 V-cycle.
    'fp32_refine': FP64 iterative refinement (FP64 residual and solution)
    around an inner CG with FP32 vectors, FP64 inner products and the FP32
    V-cycle; `cycles` counts the inner iterations."""

    def __init__(self, operator: PoissonOperator, *, multigrid: MultigridOptions | None = None,
                 precision: str = "fp64", device: int = 0, options: LinearSolverOptions | None = None, **kw):
        if options is not None:
            precision = options.preconditioner_options.get("precision", precision)
            device = options.preconditioner_options.get("device", device)
            if options.preconditioner not in (":auto", ":geometric", ":tree_geometric"):
                raise ValueError("this solver provides the tree geometric multigrid preconditioner only")
        if multigrid is None:
            multigrid = MultigridOptions()
        over = {k: kw.pop(k) for k in ("coarse_size", "pre_sweeps", "post_sweeps", "max_levels") if k in kw}
        if kw:
            raise TypeError(f"unknown keywords {sorted(kw)}")
        if over:
            multigrid = MultigridOptions(**{**vars(multigrid), **over})
        if multigrid.pre_sweeps != multigrid.post_sweeps:
            raise ValueError("CG requires equal pre- and post-smoothing sweep counts for a geometric V-cycle")
        prec = {"fp64": L.TPAMR_FP64, "fp32_vcycle": L.TPAMR_FP32_VCYCLE, "fp32": L.TPAMR_FP32_VCYCLE,
                "fp32_refine": L.TPAMR_FP32_REFINE}
        if precision not in prec:
            raise ValueError("precision must be 'fp64', 'fp32_vcycle' or 'fp32_refine'")
        self.operator = operator
        self.multigrid = multigrid
        self.precision = prec[precision]
        self.ctx = L.Context(device, self.precision)
        self.stats = LinearSolverStats()
        self._coeff_version = None
        operator.mesh.attach(self.ctx)
        opts = multigrid._c()
        L.check(L.lib().tpamr_set_multigrid(self.ctx.h, C.byref(opts)), self.ctx.h)
        self.refresh()

    @property
    def n(self):
        return self.operator.mesh.ncells

    def refresh(self):
        """Upload the operator's current face coefficients (refreshes every
        level), with its open sides and boundary coefficients when they
        changed."""
        op = self.operator
        bstate = (op.boundary_sides, getattr(op, "boundary_version", 0))
        if bstate != getattr(self, "_boundary_state", ((False,) * 4, 0)):
            sides = np.asarray([int(v) for v in op.boundary_sides], np.int32)
            kb = op.boundary_face_coefficient
            L.check(L.lib().tpamr_set_pressure_boundary(self.ctx.h, L.ptr(sides), None if kb is None else L.ptr(kb)),
                    self.ctx.h)
            self._boundary_state = bstate
        k = op.face_coefficient
        L.check(L.lib().tpamr_set_face_coefficients(self.ctx.h, L.ptr(k)), self.ctx.h)
        self._coeff_version = op.version

    def _sync(self):
        # O(1) per call: coefficient replacements bump the operator's version
        if self._coeff_version != self.operator.version:
            self.refresh()

    @property
    def levels(self):
        return [self.level_info(l) for l in range(L.lib().tpamr_num_levels(self.ctx.h))]

    def level_info(self, level):
        self._sync()
        ints = np.zeros(6, np.int32)
        dbl = np.zeros(2)
        L.check(L.lib().tpamr_level_info(self.ctx.h, level, L.ptr(ints), L.ptr(dbl)), self.ctx.h)
        return dict(ncells=int(ints[0]), nfaces=int(ints[1]), nblocks=int(ints[2]), irregular_rows=int(ints[3]),
                    irregular_nnz=int(ints[4]), max_leaf_level=int(ints[5]), omega=float(dbl[0]),
                    omega_p=float(dbl[1]))

    def level_components(self, level):
        """Level `level`'s ungauged components: per cell its 1-based slot (0:
        gauged) and per slot its root (1-based smallest cell)."""
        self._sync()
        return _level_components(self.ctx, level, self.level_info(level)["ncells"])

    def level_boundary(self, level):
        """Level `level`'s boundary face coefficients (restricted from the
        finest level's) and flags: singular, coarse_augmented (the coarsest
        solve carries the volume constraint), open_faces."""
        self._sync()
        fl = np.zeros(4, np.int32)
        L.check(L.lib().tpamr_level_boundary(self.ctx.h, level, None, L.ptr(fl)), self.ctx.h)
        kb = np.zeros(max(int(fl[3]), 1))
        L.check(L.lib().tpamr_level_boundary(self.ctx.h, level, L.ptr(kb), L.ptr(fl)), self.ctx.h)
        return dict(coefficient=kb[:int(fl[3])], singular=bool(fl[0]), coarse_augmented=bool(fl[1]),
                    open_faces=int(fl[2]))

    def apply(self, x):
        """Integrated operator A x (FactorizedEllipticOperator mul!)."""
        self._sync()
        x = L.f64(x, self.n, "x")
        y = np.zeros(self.n)
        L.check(L.lib().tpamr_apply(self.ctx.h, L.ptr(x), L.ptr(y)), self.ctx.h)
        return y

    def ldiv(self, r):
        """One V-cycle z = B r of the tree geometric multigrid (ldiv!)."""
        self._sync()
        r = L.f64(r, self.n, "r")
        z = np.zeros(self.n)
        L.check(L.lib().tpamr_vcycle(self.ctx.h, L.ptr(r), L.ptr(z)), self.ctx.h)
        return z

    def solve(self, solution, rhs, *, rtol=1.0e-10, atol=0.0, max_cycles=100, compatibility="project"):
        """solve!(solution, solver, rhs; ...) in place on `solution` (strong form)."""
        self._sync()
        if not (0 <= rtol < 1):
            raise ValueError("rtol must lie in [0, 1)")
        if atol < 0:
            raise ValueError("atol must be nonnegative")
        if max_cycles < 0:
            raise ValueError("max_cycles must be nonnegative")
        if compatibility not in ("project", "error"):
            raise ValueError("compatibility must be :project or :error")
        rhs = L.f64(rhs, self.n, "rhs")
        if solution.dtype != np.float64 or solution.shape != (self.n,) or not solution.flags.c_contiguous:
            raise L.DimensionMismatch("solution must be a contiguous float64 vector of one value per cell")
        st = L.Stats()
        L.check(L.lib().tpamr_solve(self.ctx.h, L.ptr(rhs), L.ptr(solution), float(rtol), float(atol),
                                    int(max_cycles), 1 if compatibility == "error" else 0, C.byref(st)),
                self.ctx.h)
        self.stats = LinearSolverStats._from(st, self.precision, self.ctx.h)
        return self.stats

    def run_krylov(self, rhs, x0, target, max_iterations=0):
        """One Krylov run at the plugin seam (`_run_krylov!`,
        elliptic_solver.jl:393-452): CG on the integrated, compatible `rhs`
        from `x0`, preconditioned by one V-cycle, until the strong residual
        max |r|/V <= target. Returns (x, iterations, reason)."""
        self._sync()
        rhs = L.f64(rhs, self.n, "rhs")
        x0 = L.f64(x0, self.n, "x0")
        x = np.zeros(self.n)
        it = C.c_int(0)
        reason = C.c_int(0)
        L.check(L.lib().tpamr_krylov(self.ctx.h, L.ptr(rhs), L.ptr(x0), float(target), int(max_iterations),
                                     L.ptr(x), C.byref(it), C.byref(reason)), self.ctx.h)
        return x, it.value, REASONS.get(reason.value, str(reason.value))

    def solve_device(self, d_rhs, d_x, *, rtol=1.0e-10, atol=0.0, max_cycles=100, compatibility="project"):
        """solve! on device buffers (raw device pointers, e.g. torch .data_ptr()
        of float64 tensors on the solver's GPU). The solve waits for the work
        queued on torch's current stream of that GPU."""
        self._sync()
        L.order_after_torch(self.ctx.h, self.ctx.device)
        st = L.Stats()
        L.check(L.lib().tpamr_solve_device(self.ctx.h, C.c_void_p(d_rhs), C.c_void_p(d_x), float(rtol),
                                           float(atol), int(max_cycles), 1 if compatibility == "error" else 0,
                                           C.byref(st)), self.ctx.h)
        self.stats = LinearSolverStats._from(st, self.precision, self.ctx.h)
        return self.stats

    def time_smoother(self, reps=50):
        ms = C.c_double()
        by = C.c_double()
        L.check(L.lib().tpamr_time_smoother(self.ctx.h, int(reps), C.byref(ms), C.byref(by)), self.ctx.h)
        return ms.value, by.value

    def close(self):
        self.ctx.close()


def solve(solution, solver: EllipticSolver, rhs, **kw):
    """Functional form of solve!: returns the LinearSolverStats."""
    return solver.solve(solution, rhs, **kw)


def ldiv(solver: EllipticSolver, r):
    return solver.ldiv(r)
