"""Projection steps with pressure checkpoints, and their adjoint.

The reference advances `simulate!` one step at a time; the part of a
one-velocity step on this path is the force substep followed by the
projection (`_apply_phase_forces!`, src/two_phase/stepping.jl:181-201, then
`project!`, src/navier_stokes/projection.jl:189-260), with the pressure
operator refreshed from the step's interface first
(`_refresh_pressure_substep!`, stepping.jl:73-111). `ProjectionRun` strings
such steps together on the GPU:

    refresh the operator with the step's face coefficients k_n
    u*      = u_n + dt sqrt(k_n) K (sqrt(k_n) j_n)     (the step's force jumps j_n)
    u_{n+1} = u* - dt G_n phi_n,  A_n phi_n = V b_n,  b_n = -div(u*) / dt

and keeps phi_n of every step (the checkpoints). `backward` runs the steps
in reverse with the adjoint solves on the same GPU multigrid and returns
the gradient of J(u_N) with respect to the first step's pressure rhs b_0,
the per-step face coefficients k_n and the initial velocity u_0. The
time-stepping adjoint has no reference counterpart.

`kinetic_energy` is the one-velocity KineticEnergy diagnostic
(src/simulation/diagnostics/interface.jl:128-140, 122-126), with face dual
volumes `_face_dual_volumes` (src/mesh/topology/mesh.jl:309-321).
"""
from __future__ import annotations

import time

import numpy as np

from .mesh import BlockTreeMesh
from .physics import divergence, project, project_adjoint, weighted_gradient, weighted_jump_force
from .solver import EllipticSolver


def face_dual_volumes(mesh: BlockTreeMesh):
    """|f| (h_minus + h_plus) / 2 along the face's axis."""
    ax = mesh.face_axis
    hm = mesh.block_spacing[mesh.face_minus // mesh.cells_per_block, ax]
    hp = mesh.block_spacing[mesh.face_plus // mesh.cells_per_block, ax]
    return mesh.face_area * (hm + hp) / 2


def staggered_liquid_fraction(mesh: BlockTreeMesh, fraction):
    """`_staggered_liquid_fraction` (src/two_phase/forces.jl:23-36):
    clamp((h- c- + h+ c+) / (h- + h+), 0, 1) along the face axis."""
    ax = mesh.face_axis
    m, p = mesh.face_minus, mesh.face_plus
    hm = mesh.block_spacing[m // mesh.cells_per_block, ax]
    hp = mesh.block_spacing[p // mesh.cells_per_block, ax]
    c = np.asarray(fraction, np.float64)
    return np.clip((hm * c[m] + hp * c[p]) / (hm + hp), 0.0, 1.0)


def kinetic_energy(mesh: BlockTreeMesh, u, fraction, density):
    """sum_f rho_f u_f^2 V_f / 2 with the staggered mixture density; returns
    (J, dJ/du)."""
    cf = staggered_liquid_fraction(mesh, fraction)
    rho = density[0] * cf + density[1] * (1.0 - cf)
    w = rho * face_dual_volumes(mesh)
    u = np.asarray(u, np.float64)
    return float(np.sum(w * u * u) / 2), w * u


class ProjectionRun:
    """Checkpointed force-and-projection steps on one solver (see the module
    docstring). `steps` is a sequence of (face coefficients, force jumps or
    None), one pair per step."""

    def __init__(self, solver: EllipticSolver, dt, *, tolerance=1e-12, rtol=None, max_cycles=200,
                 adjoint_rtol=1e-12):
        """`tolerance` is project!'s (absolute, on the dt^2-scaled strong
        residual); with `rtol` each step's tolerance is instead rtol times
        that step's initial strong residual max |div u*| dt."""
        if not dt > 0:
            raise ValueError("dt must be strictly positive")
        self.solver = solver
        self.dt = float(dt)
        self.tolerance = float(tolerance)
        self.rtol = None if rtol is None else float(rtol)
        self.timings = []
        self.max_cycles = int(max_cycles)
        self.adjoint_rtol = float(adjoint_rtol)
        self.steps = None
        self.checkpoints = []
        self.stats = []

    def _load(self, kappa):
        op = self.solver.operator
        op.replace_face_coefficient(kappa)
        self.solver.refresh()

    def forward(self, u0, steps, rhs_shift=None):
        """Run the steps from u0; returns u_N. `rhs_shift` (cells, strong)
        is added to the first step's pressure rhs (the projection is linear
        in it: its share of the velocity correction comes from one more
        solve)."""
        s = self.solver
        m = s.operator.mesh
        u = np.array(u0, np.float64)
        if u.shape != (m.nfaces,):
            raise ValueError(f"u0 must have length {m.nfaces}")
        self.steps = list(steps)
        self.checkpoints = []
        self.stats = []
        self.timings = []
        for n, (kappa, jumps) in enumerate(self.steps):
            t0 = time.perf_counter()
            self._load(kappa)
            t1 = time.perf_counter()
            if jumps is not None:
                u += self.dt * weighted_jump_force(s, jumps)
            tol = self.tolerance
            if self.rtol is not None:
                tol = self.rtol * float(np.max(np.abs(divergence(s, u)))) * self.dt
                tol = min(max(tol, 1e-300), 0.5)
            t2 = time.perf_counter()
            p = np.zeros(m.ncells)
            st, phi = project(u, p, s, self.dt, tolerance=tol, max_cycles=self.max_cycles)
            self.timings.append(dict(refresh_s=t1 - t0, force_s=t2 - t1, project_s=time.perf_counter() - t2))
            if n == 0 and rhs_shift is not None:
                q = np.zeros(m.ncells)
                s.solve(q, np.asarray(rhs_shift, np.float64), rtol=self.tolerance, max_cycles=self.max_cycles)
                u -= self.dt * weighted_gradient(s, q)
                phi = phi + q
            self.checkpoints.append(phi)
            self.stats.append(st)
        return u

    def backward(self, ubar):
        """Reverse sweep for dJ/du_N = ubar; returns dict(rhs = dJ/db_0,
        kappa = [dJ/dk_n], u0 = dJ/du_0, stats = adjoint solve stats)."""
        if self.steps is None:
            raise RuntimeError("forward() must run before backward()")
        s = self.solver
        ub = np.array(ubar, np.float64)
        gk = [None] * len(self.steps)
        stats = [None] * len(self.steps)
        db = None
        for n in range(len(self.steps) - 1, -1, -1):
            kappa, jumps = self.steps[n]
            self._load(kappa)
            ub, db, gk[n], stats[n] = project_adjoint(s, ub, self.dt, phi=self.checkpoints[n], jumps=jumps,
                                                      rtol=self.adjoint_rtol, max_cycles=self.max_cycles)
        return dict(rhs=db, kappa=gk, u0=ub, stats=stats)
