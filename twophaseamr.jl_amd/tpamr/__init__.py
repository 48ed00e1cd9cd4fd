"""tpamr: MI355X-native pressure-projection solver for TwoPhaseAMR.jl's tree
geometric multigrid (host mirror of the reference interface over the C ABI of
libtpamr_hip.so)."""
from ._lib import LIB_PATH, TPAMR_FP32_REFINE, TPAMR_FP32_VCYCLE, TPAMR_FP64, Context, DimensionMismatch, lib
from .hierarchy import HostHierarchy
from .mesh import BlockKey, BlockTreeMesh
from .physics import (CGFM, CSF, divergence, fill_halo, halo_schedule, mixture_face_coefficients, pressure_solve,
                      pressure_solve_rrule, project, refresh_two_phase, surface_tension, surface_tension_jumps,
                      weighted_gradient, weighted_jump_force)
from .solver import (EllipticSolver, LinearSolverOptions, LinearSolverStats, MultigridOptions, PoissonOperator, ldiv,
                     solve)

__all__ = [
    "BlockKey", "BlockTreeMesh", "CGFM", "CSF", "Context", "DimensionMismatch", "EllipticSolver", "HostHierarchy",
    "LIB_PATH", "LinearSolverOptions", "LinearSolverStats", "MultigridOptions", "PoissonOperator",
    "TPAMR_FP32_REFINE", "TPAMR_FP32_VCYCLE", "TPAMR_FP64", "divergence", "fill_halo", "halo_schedule", "ldiv", "lib", "mixture_face_coefficients",
    "pressure_solve", "pressure_solve_rrule", "project", "refresh_two_phase", "solve", "surface_tension",
    "surface_tension_jumps", "weighted_gradient", "weighted_jump_force",
]
