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
from .solver import EllipticSolver, LinearSolverStats


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


def kinetic_energy_weights(mesh: BlockTreeMesh, fraction, density):
    """rho_f V_f per face: the staggered mixture density times the face dual
    volume."""
    cf = staggered_liquid_fraction(mesh, fraction)
    rho = density[0] * cf + density[1] * (1.0 - cf)
    return rho * face_dual_volumes(mesh)


def kinetic_energy(mesh: BlockTreeMesh, u, fraction, density):
    """sum_f rho_f u_f^2 V_f / 2 with the staggered mixture density; returns
    (J, dJ/du)."""
    w = kinetic_energy_weights(mesh, fraction, density)
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


class DeviceProjectionRun:
    """ProjectionRun with every field in HBM: the velocity, the steps' face
    coefficients and force jumps, the checkpoints and the adjoint vectors are
    torch tensors (float64) on the solver's device, and each step calls the
    device-memory entry points (coefficient refresh, force, projection, its
    pullback), so nothing crosses PCIe inside the loop. `steps` is a sequence
    of (face coefficients, force jumps or None) tensors."""

    def __init__(self, solver: EllipticSolver, dt, *, tolerance=1e-12, rtol=None, max_cycles=200,
                 adjoint_rtol=1e-12):
        if not dt > 0:
            raise ValueError("dt must be strictly positive")
        self.solver = solver
        self.dt = float(dt)
        self.tolerance = float(tolerance)
        self.rtol = None if rtol is None else float(rtol)
        self.max_cycles = int(max_cycles)
        self.adjoint_rtol = float(adjoint_rtol)
        self.steps = None
        self.checkpoints = []
        self.stats = []

    def _check(self, rc):
        from . import _lib as L
        L.check(rc, self.solver.ctx.h)

    @staticmethod
    def _ptr(t):
        return None if t is None else t.data_ptr()

    def _tensor(self, t, n, what):
        import torch
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.shape == (n,)
                and t.is_contiguous()):
            raise ValueError(f"{what} must be a contiguous float64 device tensor of length {n}")
        return t

    def forward(self, u0, steps):
        import ctypes as C
        import torch
        from . import _lib as L
        s = self.solver
        s._sync()
        m = s.operator.mesh
        lib = L.lib()
        h = s.ctx.h
        u = self._tensor(u0, m.nfaces, "u0").clone()
        L.order_after_torch(h, u.device)
        div = torch.empty(m.ncells, dtype=torch.float64, device=u.device) if self.rtol is not None else None
        self.steps = [(self._tensor(k, m.nfaces, "face coefficients"),
                       None if j is None else self._tensor(j, m.nfaces, "force jumps")) for k, j in steps]
        self.checkpoints = []
        self.stats = []
        for kappa, jumps in self.steps:
            self._check(lib.tpamr_set_face_coefficients_device(h, kappa.data_ptr()))
            if jumps is not None:
                self._check(lib.tpamr_apply_force_device(h, jumps.data_ptr(), self.dt, u.data_ptr()))
            tol = self.tolerance
            if self.rtol is not None:
                self._check(lib.tpamr_divergence_device(h, u.data_ptr(), div.data_ptr()))
                tol = min(max(self.rtol * float(div.abs().max()) * self.dt, 1e-300), 0.5)
            phi = torch.empty(m.ncells, dtype=torch.float64, device=u.device)
            st = L.Stats()
            self._check(lib.tpamr_project_device(h, u.data_ptr(), None, self.dt, tol, self.max_cycles, None, self.dt,
                                                 phi.data_ptr(), C.byref(st)))
            self.checkpoints.append(phi)
            self.stats.append(LinearSolverStats._from(st, s.precision))
        return u

    def forward_interface(self, u0, interfaces, *, density, gravity):
        """`forward` with each step's face coefficients and gravity jumps
        computed on the GPU from that step's interface data: `interfaces` is a
        sequence of (cell fraction, liquid face aperture, gas face aperture,
        centroid offsets or None) device tensors; the coefficients are the
        mixture's (`mixture_face_coefficients_device`) and the jumps
        gravity!'s (`gravity_jumps_device`). They are kept as the run's steps
        for `backward`."""
        from .physics import gravity_jumps_device, mixture_face_coefficients_device
        s = self.solver
        s._sync()
        steps = []
        for fraction, liquid, gas, offsets in interfaces:
            kappa = mixture_face_coefficients_device(s, fraction, density)
            jumps = gravity_jumps_device(s, liquid, gas, density=density, gravity=gravity, centroid_offsets=offsets)
            steps.append((kappa, jumps))
        return self.forward(u0, steps)

    def backward(self, ubar):
        import ctypes as C
        import torch
        from . import _lib as L
        if self.steps is None:
            raise RuntimeError("forward() must run before backward()")
        s = self.solver
        m = s.operator.mesh
        lib = L.lib()
        h = s.ctx.h
        ub = self._tensor(ubar, m.nfaces, "ubar").clone()
        L.order_after_torch(h, ub.device)
        gk = [None] * len(self.steps)
        stats = [None] * len(self.steps)
        db = torch.empty(m.ncells, dtype=torch.float64, device=ub.device)
        for n in range(len(self.steps) - 1, -1, -1):
            kappa, jumps = self.steps[n]
            self._check(lib.tpamr_set_face_coefficients_device(h, kappa.data_ptr()))
            us = torch.empty_like(ub)
            gk[n] = torch.empty_like(ub)
            st = L.Stats()
            self._check(lib.tpamr_project_adjoint_device(h, ub.data_ptr(), self.checkpoints[n].data_ptr(),
                                                         self._ptr(jumps), self.dt, self.adjoint_rtol,
                                                         self.max_cycles, us.data_ptr(),
                                                         db.data_ptr() if n == 0 else None, gk[n].data_ptr(),
                                                         C.byref(st)))
            stats[n] = LinearSolverStats._from(st, s.precision)
            ub = us
        return dict(rhs=db, kappa=gk, u0=ub, stats=stats)
