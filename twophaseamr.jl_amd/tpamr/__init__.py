"""tpamr: MI355X-native pressure-projection solver for TwoPhaseAMR.jl's tree
geometric multigrid (host mirror of the reference interface over the C ABI of
libtpamr_hip.so)."""
from ._lib import LIB_PATH, TPAMR_FP32_REFINE, TPAMR_FP32_VCYCLE, TPAMR_FP64, Context, DimensionMismatch, lib
from .hierarchy import HostHierarchy
from .mesh import BlockKey, BlockTreeMesh
from .physics import (CGFM, CSF, divergence, fill_halo, gravity, gravity_jumps, gravity_jumps_device, halo_schedule,
                      mixture_boundary_coefficients, mixture_face_coefficients, mixture_face_coefficients_device,
                      pressure_predictor, pressure_solve, pressure_solve_rrule, project,
                      project_adjoint, project_open, weighted_gradient_boundary,
                      refresh_two_phase, surface_tension, surface_tension_jumps, surface_tension_jumps_from_cells,
                      weighted_gradient,
                      weighted_jump_force)
from .cmdgfm import CMDGFMBlockSystem, CMDGFMSolveStats
from .navier_stokes import NavierStokesSystem, step, step_device
from .stepping import DeviceProjectionRun, ProjectionRun, face_dual_volumes, kinetic_energy, staggered_liquid_fraction
from .solver import (EllipticSolver, FreeSlip, LinearSolverOptions, LinearSolverStats, MultigridOptions, NoSlip,
                     Outflow, PoissonOperator, VelocityBoundaryConditions, ZeroGradientVelocity, ldiv,
                     pressure_boundary_sides, solve)

__all__ = [
    "BlockKey", "BlockTreeMesh", "CGFM", "CMDGFMBlockSystem", "CMDGFMSolveStats", "CSF", "Context",
    "DeviceProjectionRun", "DimensionMismatch", "EllipticSolver",
    "FreeSlip", "NoSlip", "Outflow", "VelocityBoundaryConditions", "ZeroGradientVelocity", "pressure_boundary_sides",
    "project_open", "weighted_gradient_boundary",
    "HostHierarchy",
    "LIB_PATH", "LinearSolverOptions", "LinearSolverStats", "MultigridOptions", "NavierStokesSystem",
    "PoissonOperator", "ProjectionRun", "step", "step_device",
    "TPAMR_FP32_REFINE", "TPAMR_FP32_VCYCLE", "TPAMR_FP64", "divergence", "face_dual_volumes", "fill_halo",
    "gravity", "gravity_jumps", "gravity_jumps_device", "halo_schedule", "kinetic_energy", "ldiv", "lib",
    "mixture_boundary_coefficients", "mixture_face_coefficients", "mixture_face_coefficients_device",
    "pressure_predictor", "pressure_solve", "pressure_solve_rrule", "project", "project_adjoint", "refresh_two_phase", "solve",
    "staggered_liquid_fraction", "surface_tension", "surface_tension_jumps", "surface_tension_jumps_from_cells",
    "weighted_gradient",
    "weighted_jump_force",
]
