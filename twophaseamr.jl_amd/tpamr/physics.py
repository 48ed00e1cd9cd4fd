"""Two-phase pressure coefficients, surface-tension jumps, projection and the
pressure adjoint, on the GPU.

Reference interfaces (TwoPhaseAMR.jl):
  mixture_density / PoissonOperator(::InterfaceField; density)  src/two_phase/forces.jl:9-140
  _surface_tension_jumps! (CGFM / CSF), _face_curvature         src/two_phase/forces.jl:211-378
  _weighted_jump_force!                                         src/two_phase/forces.jl:179-209
  divergence!                                                   src/operators/advection/cell_advection.jl:205-220
  weighted_gradient!                                            src/operators/elliptic.jl:1272-1297
  project! / _project_prepared! / _projection_rhs!              src/navier_stokes/projection.jl:100-260
  _pressure_predictor!                                          src/single_phase/stepping.jl:144-168
  fill_halo! / _build_halo_schedule                             src/mesh/variables.jl:205-229, mesh.jl:944-1037
  gravity!                                                      src/two_phase/forces.jl:528-580
The adjoints (`pressure_solve_rrule`, `project_adjoint`) have no reference
counterpart.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .solver import EllipticSolver, LinearSolverStats, PoissonOperator

CGFM = 0
CSF = 1


def mixture_face_coefficients(solver: EllipticSolver, fraction, density=(1.0, 1.0)):
    """inv.(mixture_density(field, density)): kappa_f = 1 / (rho_l c_f + rho_g (1 - c_f))."""
    rho_l, rho_g = (float(v) for v in density)
    if not (np.isfinite(rho_l) and rho_l > 0 and np.isfinite(rho_g) and rho_g > 0):
        raise ValueError("phase densities must be finite and positive")
    m = solver.operator.mesh
    c = L.f64(fraction, m.ncells, "fraction")
    k = np.zeros(m.nfaces)
    L.check(L.lib().tpamr_mixture_face_coefficients(solver.ctx.h, L.ptr(c), rho_l, rho_g, L.ptr(k)), solver.ctx.h)
    return k


def mixture_boundary_coefficients(solver: EllipticSolver, fraction, density=(1.0, 1.0)):
    """`_mixture_boundary_coefficient` (forces.jl:103-111): per boundary face
    1 / (rho_l c + rho_g (1 - c)) of its cell's fraction."""
    rho_l, rho_g = (float(v) for v in density)
    m = solver.operator.mesh
    c = L.f64(fraction, m.ncells, "fraction")
    kb = np.zeros(max(m.nboundary_faces, 1))
    L.check(L.lib().tpamr_mixture_boundary_coefficients(solver.ctx.h, L.ptr(c), rho_l, rho_g, L.ptr(kb)),
            solver.ctx.h)
    return kb[:m.nboundary_faces]


def refresh_two_phase(solver: EllipticSolver, fraction, density):
    """`_refresh_pressure_substep!` (src/two_phase/stepping.jl:73-111) with
    `PoissonOperator(field; density, template)` (forces.jl:112-140): new
    mixture coefficients into the operator and every multigrid level, and,
    with open sides, the boundary faces' mixture coefficients."""
    k = mixture_face_coefficients(solver, fraction, density)
    kb = mixture_boundary_coefficients(solver, fraction, density) if any(solver.operator.boundary_sides) else None
    solver.operator.replace_face_coefficient(k, boundary_face_coefficient=kb)
    solver.refresh()
    return solver


def surface_tension_jumps(solver: EllipticSolver, fraction, face_curvature, *, sigma, formulation=CGFM):
    """Integrated jumps sigma kappa_f [H]_f |f| (NaN curvature = unresolved face)."""
    if not (np.isfinite(sigma) and sigma >= 0):
        raise ValueError("sigma must be finite and nonnegative")
    if formulation not in (CGFM, CSF):
        raise ValueError("formulation must be CSF() or CGFM()")
    m = solver.operator.mesh
    c = L.f64(fraction, m.ncells, "fraction")
    curv = np.ascontiguousarray(np.broadcast_to(np.asarray(face_curvature, np.float64), (m.nfaces,)))
    out = np.zeros(m.nfaces)
    L.check(L.lib().tpamr_surface_tension_jumps(solver.ctx.h, L.ptr(c), L.ptr(curv), float(sigma), int(formulation),
                                                L.ptr(out)), solver.ctx.h)
    return out


def surface_tension_jumps_from_cells(solver: EllipticSolver, fraction, cell_curvature, curvature_quality, *, sigma,
                                     formulation=CGFM):
    """`_surface_tension_jumps!` (src/two_phase/forces.jl:345-378) with the
    face curvature of `_face_curvature` (forces.jl:211-300) from the
    reconstruction's cell curvatures (NaN = none) and their height-function
    coverage. Returns (jumps, number of unresolved faces)."""
    if not (np.isfinite(sigma) and sigma >= 0):
        raise ValueError("sigma must be finite and nonnegative")
    if formulation not in (CGFM, CSF):
        raise ValueError("formulation must be CSF() or CGFM()")
    m = solver.operator.mesh
    c = L.f64(fraction, m.ncells, "fraction")
    k = L.f64(cell_curvature, m.ncells, "cell curvature")
    q = L.f64(curvature_quality, m.ncells, "curvature quality")
    out = np.zeros(m.nfaces)
    n = C.c_int()
    L.check(L.lib().tpamr_surface_tension_jumps_cells(solver.ctx.h, L.ptr(c), L.ptr(k), L.ptr(q), float(sigma),
                                                      int(formulation), L.ptr(out), C.byref(n)), solver.ctx.h)
    return out, n.value


def weighted_jump_force(solver: EllipticSolver, jumps):
    solver._sync()
    m = solver.operator.mesh
    j = L.f64(jumps, m.nfaces, "jumps")
    out = np.zeros(m.nfaces)
    L.check(L.lib().tpamr_weighted_jump_force(solver.ctx.h, L.ptr(j), L.ptr(out)), solver.ctx.h)
    return out


def surface_tension(solver: EllipticSolver, fraction, face_curvature, *, sigma, formulation=CGFM):
    """surface_tension! interior face accelerations (boundary faces are zero)."""
    return weighted_jump_force(solver, surface_tension_jumps(solver, fraction, face_curvature, sigma=sigma,
                                                             formulation=formulation))


def _gravity_args(solver, liquid_aperture, gas_aperture, density, gravity, centroid_offsets):
    m = solver.operator.mesh
    rho_l, rho_g = (float(v) for v in density)
    gx, gy = (float(v) for v in gravity)
    la = L.f64(liquid_aperture, m.nfaces, "liquid aperture")
    ga = L.f64(gas_aperture, m.nfaces, "gas aperture")
    off = None
    if centroid_offsets is not None:
        off = np.ascontiguousarray(centroid_offsets, np.float64)
        if off.shape not in ((m.ncells, 2), (2 * m.ncells,)):
            raise L.DimensionMismatch(f"centroid offsets must have shape ({m.ncells}, 2)")
    return m, (L.ptr(off), L.ptr(la), L.ptr(ga), rho_l, rho_g, gx, gy), (off, la, ga)


def gravity_jumps(solver: EllipticSolver, liquid_aperture, gas_aperture, *, density, gravity, centroid_offsets=None):
    """The face jumps of gravity! (src/two_phase/forces.jl:528-580):
    rho_f |f| (phi+ - phi-) with phi = g . x at the cell centres, moved in mixed
    cells by `centroid_offsets` (per cell, the interface centroid relative to
    the cell centre in units of the cell spacing; None: cell centres) and
    rho_f = rho_l a_l + rho_g a_g from the phase face apertures."""
    m, args, _keep = _gravity_args(solver, liquid_aperture, gas_aperture, density, gravity, centroid_offsets)
    out = np.zeros(m.nfaces)
    L.check(L.lib().tpamr_gravity_jumps(solver.ctx.h, *args, L.ptr(out)), solver.ctx.h)
    return out


def gravity(solver: EllipticSolver, liquid_aperture, gas_aperture, *, density, gravity, centroid_offsets=None):
    """gravity! interior face accelerations: the jumps of `gravity_jumps`
    through `_weighted_jump_force!`."""
    solver._sync()
    m, args, _keep = _gravity_args(solver, liquid_aperture, gas_aperture, density, gravity, centroid_offsets)
    out = np.zeros(m.nfaces)
    L.check(L.lib().tpamr_gravity(solver.ctx.h, *args, L.ptr(out)), solver.ctx.h)
    return out


def _device_vec(t, n, what, width=1):
    import torch
    shape = (n,) if width == 1 else (n, width)
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
            and (tuple(t.shape) == shape or (width > 1 and tuple(t.shape) == (n * width,)))):
        raise ValueError(f"{what} must be a contiguous float64 device tensor of shape {shape}")
    return t


def mixture_face_coefficients_device(solver: EllipticSolver, fraction, density=(1.0, 1.0), out=None):
    """`mixture_face_coefficients` on device memory: `fraction` (cells) and the
    result (faces, a new tensor unless `out` is given) are float64 tensors on
    the solver's GPU."""
    import torch
    rho_l, rho_g = (float(v) for v in density)
    m = solver.operator.mesh
    c = _device_vec(fraction, m.ncells, "fraction")
    k = torch.empty(m.nfaces, dtype=torch.float64, device=c.device) if out is None else \
        _device_vec(out, m.nfaces, "out")
    L.order_after_torch(solver.ctx.h, c.device)
    L.check(L.lib().tpamr_mixture_face_coefficients_device(solver.ctx.h, c.data_ptr(), rho_l, rho_g, k.data_ptr()),
            solver.ctx.h)
    return k


def gravity_jumps_device(solver: EllipticSolver, liquid_aperture, gas_aperture, *, density, gravity,
                         centroid_offsets=None, out=None):
    """`gravity_jumps` on device memory: apertures (faces), offsets (cells x 2
    or None) and the result are float64 tensors on the solver's GPU."""
    import torch
    m = solver.operator.mesh
    rho_l, rho_g = (float(v) for v in density)
    gx, gy = (float(v) for v in gravity)
    la = _device_vec(liquid_aperture, m.nfaces, "liquid aperture")
    ga = _device_vec(gas_aperture, m.nfaces, "gas aperture")
    off = None if centroid_offsets is None else _device_vec(centroid_offsets, m.ncells, "centroid offsets", 2)
    j = torch.empty(m.nfaces, dtype=torch.float64, device=la.device) if out is None else \
        _device_vec(out, m.nfaces, "out")
    L.order_after_torch(solver.ctx.h, la.device)
    L.check(L.lib().tpamr_gravity_jumps_device(solver.ctx.h, None if off is None else off.data_ptr(), la.data_ptr(),
                                               ga.data_ptr(), rho_l, rho_g, gx, gy, j.data_ptr()), solver.ctx.h)
    return j


def divergence(solver: EllipticSolver, u, u_boundary=None):
    m = solver.operator.mesh
    uu = L.f64(u, m.nfaces, "u")
    ub = None if u_boundary is None else L.f64(u_boundary, m.nboundary_faces, "u_boundary")
    out = np.zeros(m.ncells)
    L.check(L.lib().tpamr_divergence(solver.ctx.h, L.ptr(uu), L.ptr(ub), L.ptr(out)), solver.ctx.h)
    return out


def weighted_gradient(solver: EllipticSolver, p):
    solver._sync()
    m = solver.operator.mesh
    pp = L.f64(p, m.ncells, "p")
    out = np.zeros(m.nfaces)
    L.check(L.lib().tpamr_weighted_gradient(solver.ctx.h, L.ptr(pp), L.ptr(out)), solver.ctx.h)
    return out


def weighted_gradient_boundary(solver: EllipticSolver, p):
    """`_elliptic_boundary_gradient!` (elliptic.jl:1254-1263): per boundary
    face -side (2 k_b / h) p_cell on open faces, 0 on closed ones."""
    solver._sync()
    m = solver.operator.mesh
    pp = L.f64(p, m.ncells, "p")
    out = np.zeros(max(m.nboundary_faces, 1))
    L.check(L.lib().tpamr_weighted_gradient_boundary(solver.ctx.h, L.ptr(pp), L.ptr(out)), solver.ctx.h)
    return out[:m.nboundary_faces]


def project_open(u, ub, p, solver: EllipticSolver, dt, *, tolerance=float(np.sqrt(np.finfo(float).eps)),
                 max_cycles=100):
    """project! on a domain with open pressure sides: the boundary face
    velocities `ub` enter the divergence and the open faces receive
    -dt g_b; u, ub and p are updated in place; returns (stats, phi)."""
    solver._sync()
    m = solver.operator.mesh
    for name, v, n in (("u", u, m.nfaces), ("ub", ub, m.nboundary_faces), ("p", p, m.ncells)):
        if v.dtype != np.float64 or v.shape != (n,) or not v.flags.c_contiguous:
            raise L.DimensionMismatch(f"{name} must be a contiguous float64 vector of {n} values")
    phi = np.zeros(m.ncells)
    st = L.Stats()
    L.check(L.lib().tpamr_project_open(solver.ctx.h, L.ptr(u), L.ptr(ub), L.ptr(p), float(dt), float(tolerance),
                                       int(max_cycles), L.ptr(phi), C.byref(st)), solver.ctx.h)
    return LinearSolverStats._from(st, solver.precision, solver.ctx.h), phi


def pressure_predictor(u, p, solver: EllipticSolver, dt):
    """`_pressure_predictor!` (src/single_phase/stepping.jl:144-168): u -= dt
    G p with the operator's weighted gradient, in place; returns u."""
    solver._sync()
    m = solver.operator.mesh
    if u.dtype != np.float64 or u.shape != (m.nfaces,) or not u.flags.c_contiguous:
        raise L.DimensionMismatch("u must be a contiguous float64 vector over interior faces")
    pp = L.f64(p, m.ncells, "p")
    L.check(L.lib().tpamr_pressure_predictor(solver.ctx.h, L.ptr(u), L.ptr(pp), float(dt)), solver.ctx.h)
    return u


def project(u, p, solver: EllipticSolver, dt, *, tolerance=float(np.sqrt(np.finfo(float).eps)), max_cycles=100,
            inherited_divergence=None, reference_dt=None):
    """project! on a wall/periodic domain: u and p are updated in place; returns
    (stats, phi). `inherited_divergence` and `reference_dt` are project!'s
    `_inherited_divergence` and `_reference_dt` (a shortened step keeps part of
    the inherited divergence, projection.jl:130-156)."""
    solver._sync()
    m = solver.operator.mesh
    if u.dtype != np.float64 or u.shape != (m.nfaces,) or not u.flags.c_contiguous:
        raise L.DimensionMismatch("u must be a contiguous float64 vector over interior faces")
    if p.dtype != np.float64 or p.shape != (m.ncells,) or not p.flags.c_contiguous:
        raise L.DimensionMismatch("p must be a contiguous float64 vector over cells")
    inh = None if inherited_divergence is None else L.f64(inherited_divergence, m.ncells, "inherited divergence")
    phi = np.zeros(m.ncells)
    st = L.Stats()
    L.check(L.lib().tpamr_project_ex(solver.ctx.h, L.ptr(u), L.ptr(p), float(dt), float(tolerance), int(max_cycles),
                                     L.ptr(inh), float(dt if reference_dt is None else reference_dt), L.ptr(phi),
                                     C.byref(st)), solver.ctx.h)
    return LinearSolverStats._from(st, solver.precision, solver.ctx.h), phi


def pressure_solve(solver: EllipticSolver, rhs, *, rtol=1e-12, max_cycles=200):
    p = np.zeros(solver.n)
    st = solver.solve(p, rhs, rtol=rtol, max_cycles=max_cycles)
    return p, st


def pressure_solve_rrule(solver: EllipticSolver, rhs, *, rtol=1e-12, max_cycles=200):
    """Forward pressure solve and its pullback, like a ChainRulesCore rrule.

    Returns (p, pullback) with pullback(w) -> (dJ/drhs, dJ/dkappa) for J = w' p.
    The adjoint solve reuses the same (symmetric) GPU multigrid."""
    p, _ = pressure_solve(solver, rhs, rtol=rtol, max_cycles=max_cycles)

    def pullback(w):
        m = solver.operator.mesh
        ww = L.f64(w, m.ncells, "w")
        db = np.zeros(m.ncells)
        dk = np.zeros(m.nfaces)
        st = L.Stats()
        L.check(L.lib().tpamr_adjoint(solver.ctx.h, L.ptr(ww), L.ptr(p), float(rtol), int(max_cycles), L.ptr(db),
                                      L.ptr(dk), C.byref(st)), solver.ctx.h)
        return db, dk

    return p, pullback


def project_adjoint(solver: EllipticSolver, ubar, dt, *, phi=None, jumps=None, rtol=1e-12, max_cycles=200):
    """Pullback of one forcing-plus-projection step u' = P(u + dt a), a =
    weighted_jump_force(jumps), through `project` with the current operator.

    Given ubar = dJ/du' returns (dJ/du, dJ/drhs, dJ/dkappa, stats): dJ/du is
    also dJ/du* (u* = u + dt a); dJ/drhs is the gradient with respect to the
    step's strong pressure rhs b = -div(u*) / dt; dJ/dkappa (None without
    `phi`, the step's checkpointed pressure increment) is the gradient with
    respect to the face coefficients of the step (operator, gradient and force)."""
    solver._sync()
    m = solver.operator.mesh
    ub = L.f64(ubar, m.nfaces, "ubar")
    ph = None if phi is None else L.f64(phi, m.ncells, "phi")
    jp = None if jumps is None else L.f64(jumps, m.nfaces, "jumps")
    us = np.zeros(m.nfaces)
    db = np.zeros(m.ncells)
    dk = None if phi is None else np.zeros(m.nfaces)
    st = L.Stats()
    L.check(L.lib().tpamr_project_adjoint(solver.ctx.h, L.ptr(ub), L.ptr(ph), L.ptr(jp), float(dt), float(rtol),
                                          int(max_cycles), L.ptr(us), L.ptr(db), L.ptr(dk), C.byref(st)),
            solver.ctx.h)
    return us, db, dk, LinearSolverStats._from(st, solver.precision)


def halo_schedule(ctx: L.Context, halo=1):
    """The halo refresh schedule of `_build_halo_schedule`
    (src/mesh/topology/mesh.jl:944-1037) as built by the library: per entry
    its offset in the haloed output, donor cells with weights (CSR) and its
    rank among the physical-boundary entries (-1 otherwise)."""
    ne, nphys = C.c_int(), C.c_int()
    L.check(L.lib().tpamr_halo_sizes(ctx.h, int(halo), C.byref(ne), C.byref(nphys)), ctx.h)
    dest = np.zeros(ne.value, np.int32)
    ptr = np.zeros(ne.value + 1, np.int32)
    phys = np.zeros(ne.value, np.int32)
    probe_n = max(ne.value * 4, 1)
    donor = np.zeros(probe_n, np.int32)
    weight = np.zeros(probe_n)
    L.check(L.lib().tpamr_halo_schedule(ctx.h, int(halo), L.ptr(dest), L.ptr(ptr), L.ptr(donor), L.ptr(weight),
                                        L.ptr(phys)), ctx.h)
    nd = int(ptr[-1])
    return dict(dest=dest, ptr=ptr, donor=donor[:nd], weight=weight[:nd], phys=phys, nphysical=nphys.value)


def fill_halo(solver: EllipticSolver, cells, halo=1, physical=None):
    """fill_halo!(field, ZeroGradient()) or, with `physical` (one value per
    physical halo entry in schedule order), fill_halo!(field, Dirichlet(f))
    (src/mesh/variables.jl:205-229) on the GPU: returns per block the haloed
    array, shape (nblocks, by + 2 halo, bx + 2 halo) (CellField.blocks with the
    x index fastest)."""
    m = solver.operator.mesh
    bx, by = m.block_size
    cells = L.f64(cells, m.ncells, "cell values")
    out = np.zeros((m.nblocks, by + 2 * halo, bx + 2 * halo))
    php = None
    if physical is not None:
        ne, nphys = C.c_int(), C.c_int()
        L.check(L.lib().tpamr_halo_sizes(solver.ctx.h, int(halo), C.byref(ne), C.byref(nphys)), solver.ctx.h)
        physical = L.f64(physical, nphys.value, "physical halo values")
        php = L.ptr(physical)
    L.check(L.lib().tpamr_fill_halo(solver.ctx.h, int(halo), L.ptr(cells), php, L.ptr(out)), solver.ctx.h)
    return out
