"""One-velocity Navier-Stokes step of one phase on a uniform mesh periodic in
both directions: host mirror of the reference's NavierStokesSystem,
NavierStokesWorkspace and ``step!`` over the C ABI.

Reference interface (TwoPhaseAMR.jl):
- ``NavierStokesSystem(mesh; density, dynamic_viscosity, ...)``
  (src/navier_stokes/system.jl): pressure face coefficient 1/rho, kinematic
  viscosity mu/rho, staggered viscous operator and transport options
  (``AdvectionOptions(; scheme=LaxWendroff())``, navier_stokes/options.jl:1-35);
- ``step!(velocity, pressure, solver, dt, workspace; force, tolerance,
  max_cycles)`` (src/single_phase/stepping.jl:194-264);
- ``advect!``, ``diffusion!``, ``implicit_midpoint_diffusion!``,
  ``stable_timestep`` (src/operators/advection/staggered_update.jl:265-455,
  src/operators/staggered_diffusion.jl:880-900,
  src/navier_stokes/implicit_diffusion.jl:916-965).

The staggered transport and viscous operators take the reference's uniform
periodic schedule (``_uniform_periodic_cell_faces``,
staggered_schedules.jl:226-264); other meshes raise ValueError, as that
schedule does. Velocities are interior-face vectors in the mesh's face order;
pressures are cell vectors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .mesh import BlockTreeMesh
from .solver import EllipticSolver, LinearSolverStats, MultigridOptions, PoissonOperator

SCHEMES = {"upwind": 0, "centred": 1, "centered": 1, "central": 1, "lw": 2, "lax_wendroff": 2}

_SIGS = {
    "tpamr_set_momentum": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "tpamr_momentum_sweep_forward": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "tpamr_advect": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
    "tpamr_diffusion": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tpamr_implicit_diffusion": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                           C.c_int, C.c_void_p]),
    "tpamr_stable_timestep": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_double)]),
    "tpamr_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_double, C.c_int,
                             C.POINTER(L.Stats)]),
    "tpamr_step_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_double,
                                    C.c_int, C.POINTER(L.Stats)]),
    "tpamr_step_viscous_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tpamr_momentum_levels": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_void_p]),
}


def _lib():
    lib = L.lib()
    if not getattr(lib, "_momentum_bound", False):
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib._momentum_bound = True
    return lib


def _viscous_stats(st: L.Stats) -> LinearSolverStats:
    return LinearSolverStats(cycles=st.cycles, converged=bool(st.converged), compatible=True,
                             initial_l2=st.initial_l2, initial_linf=st.initial_linf, final_l2=st.final_l2,
                             final_linf=st.final_linf, target_linf=st.target_linf,
                             residual_scale=st.residual_scale, removed_rhs_mean=0.0, method="cg",
                             preconditioner="tree_staggered_geometric_multigrid", krylov_runs=st.krylov_runs)


class NavierStokesSystem:
    """NavierStokesSystem(mesh; density=1, dynamic_viscosity, scheme='lw',
    multigrid options, device=0) of one phase: the pressure solver
    (EllipticSolver of PoissonOperator(mesh, 1/rho)) and, on its context,
    the staggered transport workspace and viscous operator with kinematic
    viscosity mu/rho. Holds the NavierStokesWorkspace state as well (the
    transport's sweep order)."""

    def __init__(self, mesh: BlockTreeMesh, *, density=1.0, dynamic_viscosity, scheme="lw",
                 multigrid: MultigridOptions | None = None, device=0, **mg):
        density = float(density)
        dynamic_viscosity = float(dynamic_viscosity)
        if not (np.isfinite(density) and density > 0):
            raise ValueError("density must be finite and positive")
        if not (np.isfinite(dynamic_viscosity) and dynamic_viscosity > 0):
            raise ValueError("dynamic_viscosity must be finite and positive")
        if scheme not in SCHEMES:
            raise ValueError(f"unsupported staggered advection scheme: {scheme}")
        self.mesh = mesh
        self.density = density
        self.dynamic_viscosity = dynamic_viscosity
        self.kinematic_viscosity = dynamic_viscosity / density
        self.scheme = scheme
        self.pressure_solver = EllipticSolver(PoissonOperator(mesh, 1.0 / density), multigrid=multigrid,
                                              precision="fp64", device=device, **mg)
        self.ctx = self.pressure_solver.ctx
        L.check(_lib().tpamr_set_momentum(self.ctx.h, self.kinematic_viscosity, SCHEMES[scheme]), self.ctx.h)

    def close(self):
        self.pressure_solver.close()

    def _face(self, v, name):
        n = self.mesh.nfaces
        if not (isinstance(v, np.ndarray) and v.dtype == np.float64 and v.shape == (n,) and v.flags.c_contiguous):
            raise L.DimensionMismatch(f"{name} must be a contiguous float64 vector over interior faces")
        return v

    @property
    def sweep_forward(self) -> bool:
        out = C.c_int()
        L.check(_lib().tpamr_momentum_sweep_forward(self.ctx.h, C.byref(out)), self.ctx.h)
        return bool(out.value)

    def advect(self, field, advecting, dt):
        """advect!(field, advecting, dt, workspace) in place; `advecting is
        field` is the one-velocity shorthand. Returns field."""
        self._face(field, "field")
        adv = field if advecting is field else L.f64(advecting, self.mesh.nfaces, "advecting velocity")
        L.check(_lib().tpamr_advect(self.ctx.h, L.ptr(field), L.ptr(adv), float(dt)), self.ctx.h)
        return field

    def diffusion(self, u):
        """diffusion!: nu L u (both components)."""
        u = L.f64(u, self.mesh.nfaces, "velocity")
        out = np.zeros(self.mesh.nfaces)
        L.check(_lib().tpamr_diffusion(self.ctx.h, L.ptr(u), L.ptr(out)), self.ctx.h)
        return out

    def implicit_midpoint_diffusion(self, dst, convected, old, dt, *,
                                    tolerance=float(np.sqrt(np.finfo(float).eps)), max_cycles=100):
        """implicit_midpoint_diffusion! in place on dst (the initial guess);
        returns (stats_x, stats_y)."""
        self._face(dst, "destination")
        cv = L.f64(convected, self.mesh.nfaces, "convected velocity")
        od = L.f64(old, self.mesh.nfaces, "old velocity")
        st = (L.Stats * 2)()
        L.check(_lib().tpamr_implicit_diffusion(self.ctx.h, L.ptr(dst), L.ptr(cv), L.ptr(od), float(dt),
                                                float(tolerance), int(max_cycles), C.cast(st, C.c_void_p)),
                self.ctx.h)
        return _viscous_stats(st[0]), _viscous_stats(st[1])

    def stable_timestep(self, u, *, cfl=0.45):
        """stable_timestep(velocity, solver, workspace; cfl): the staggered
        transport bound (cfl in (0, 1] for one phase)."""
        if not (0 < cfl <= 1):
            raise ValueError("cfl must lie in (0, 1] for one phase")
        u = L.f64(u, self.mesh.nfaces, "velocity")
        out = C.c_double()
        L.check(_lib().tpamr_stable_timestep(self.ctx.h, L.ptr(u), float(cfl), C.byref(out)), self.ctx.h)
        return out.value

    @property
    def viscous_levels(self):
        """Cells per level of the viscous solve's staggered multigrid."""
        n = C.c_int()
        cells = np.zeros(64, np.int32)
        L.check(_lib().tpamr_momentum_levels(self.ctx.h, C.byref(n), L.ptr(cells)), self.ctx.h)
        return [int(v) for v in cells[:n.value]]

    def viscous_stats(self):
        st = (L.Stats * 2)()
        L.check(_lib().tpamr_step_viscous_stats(self.ctx.h, C.cast(st, C.c_void_p)), self.ctx.h)
        return _viscous_stats(st[0]), _viscous_stats(st[1])


def step(u, p, system: NavierStokesSystem, dt, *, force=None, tolerance=float(np.sqrt(np.finfo(float).eps)),
         max_cycles=None):
    """step!(velocity, pressure, solver, dt, workspace; force, tolerance,
    max_cycles): one split momentum step and one incremental pressure
    projection, in place on u (interior faces) and p (cells). `force`: face
    values of the body force at the midpoint time. Returns the projection's
    LinearSolverStats."""
    system._face(u, "velocity")
    m = system.mesh
    if not (isinstance(p, np.ndarray) and p.dtype == np.float64 and p.shape == (m.ncells,) and p.flags.c_contiguous):
        raise L.DimensionMismatch("pressure must be a contiguous float64 vector over cells")
    f = None if force is None else L.f64(force, m.nfaces, "force")
    cycles = 100 if max_cycles is None else int(max_cycles)
    st = L.Stats()
    system.pressure_solver._sync()
    L.check(_lib().tpamr_step(system.ctx.h, L.ptr(u), L.ptr(p), float(dt), L.ptr(f), float(tolerance), cycles,
                              C.byref(st)), system.ctx.h)
    return LinearSolverStats._from(st, L.TPAMR_FP64, system.ctx.h)


def step_device(u, p, system: NavierStokesSystem, dt, *, force=None, tolerance=float(np.sqrt(np.finfo(float).eps)),
                max_cycles=None):
    """step! with u, p (and force) as torch float64 tensors in HBM on the
    system's device; updated in place."""
    import torch

    m = system.mesh
    for t, n, what in ((u, m.nfaces, "velocity"), (p, m.ncells, "pressure")):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.shape == (n,)
                and t.is_contiguous()):
            raise L.DimensionMismatch(f"{what} must be a contiguous float64 device tensor of length {n}")
    if force is not None and not (isinstance(force, torch.Tensor) and force.is_cuda and force.dtype == torch.float64
                                  and force.shape == (m.nfaces,) and force.is_contiguous()):
        raise L.DimensionMismatch("force must be a contiguous float64 device tensor over interior faces")
    cycles = 100 if max_cycles is None else int(max_cycles)
    st = L.Stats()
    system.pressure_solver._sync()
    L.order_after_torch(system.ctx.h, system.ctx.device)
    L.check(_lib().tpamr_step_device(system.ctx.h, C.c_void_p(u.data_ptr()), C.c_void_p(p.data_ptr()), float(dt),
                                     None if force is None else C.c_void_p(force.data_ptr()), float(tolerance),
                                     cycles, C.byref(st)), system.ctx.h)
    return LinearSolverStats._from(st, L.TPAMR_FP64, system.ctx.h)
