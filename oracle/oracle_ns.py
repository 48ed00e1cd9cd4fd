"""CPU FP64 restatement of the one-velocity momentum step on uniform, fully
periodic block meshes: staggered advection, the staggered viscous operator,
the implicit-midpoint viscous substep and the staged ``step!`` around the
pressure projection.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg as
the checker. The product package never imports this module.

Reference files (TwoPhaseAMR.jl):
- schedule: ``_uniform_periodic_cell_faces`` and ``_uniform_dual_schedule``
  (src/operators/advection/staggered_schedules.jl:226-264, 396-458), output
  grouping ``_dual_adjacency`` (:67-80) with ``_grouped_output_schedule``
  (src/threading.jl:455-475);
- face states: ``_dual_state`` for Upwind, Centred and LaxWendroff
  (staggered_schedules.jl:943-968);
- convection and transport: ``_directional_flux_divergence!``,
  ``_convection_schedule!``, ``_convection!``
  (src/operators/advection/staggered_fluxes.jl:180-211, 241-287, 346-381) and
  the velocity ``advect!`` (staggered_update.jl:265-326), CFL bound
  ``_accumulate_staggered_rates!``/``stable_timestep`` (:364-392, 426-455);
- viscous operator: ``_dual_diffusion_connections`` (uniform form),
  ``_minimum_norm_gradient`` (aligned pairs), ``_diffusion_stiffness`` and
  ``_staggered_diffusion_component`` (src/operators/staggered_diffusion.jl:
  332-345, 548-564, 673-712, 724-763), ``diffusion!`` (:851-900) and the
  shifted matrix ``_staggered_stable_matrix`` (src/operators/elliptic.jl:
  1344-1352);
- viscous substep ``implicit_midpoint_diffusion!``
  (src/navier_stokes/implicit_diffusion.jl:916-965);
- staged step ``step!`` (src/single_phase/stepping.jl:194-264) with the body
  force ``_add_body_force!`` (src/navier_stokes/stages.jl:40-53).

The viscous system is solved here exactly (sparse LU): the reference's Krylov
run stops within its tolerance of this solution, so parity tests run both
sides at a tight tolerance.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

SCHEMES = ("upwind", "centred", "lw")


class UniformSchedule:
    """Dual control volumes of both velocity components on a uniform, fully
    periodic mesh (``_uniform_dual_schedule``)."""

    def __init__(self, mesh):
        if mesh.nbfaces if hasattr(mesh, "nbfaces") else mesh.nboundary_faces:
            raise ValueError("staggered advection currently requires periodic boundaries in both directions")
        levels = np.asarray(mesh.block_keys)[:, 0]
        if len(levels) == 0 or np.any(levels != levels[0]):
            raise ValueError("staggered advection currently requires a uniform tree level")
        n, nf = mesh.ncells, mesh.nfaces
        fm, fp, ax = (np.asarray(a) for a in (mesh.face_minus, mesh.face_plus, mesh.face_axis))
        xminus = np.full(n, -1)
        xplus = np.full(n, -1)
        yminus = np.full(n, -1)
        yplus = np.full(n, -1)
        for f in range(nf):
            if ax[f] == 0:
                if xplus[fm[f]] >= 0 or xminus[fp[f]] >= 0:
                    raise ValueError("nonconforming x-face schedule")
                xplus[fm[f]] = f
                xminus[fp[f]] = f
            else:
                if yplus[fm[f]] >= 0 or yminus[fp[f]] >= 0:
                    raise ValueError("nonconforming y-face schedule")
                yplus[fm[f]] = f
                yminus[fp[f]] = f
        if min(xminus.min(), xplus.min(), yminus.min(), yplus.min()) < 0:
            raise ValueError("incomplete periodic face schedule")
        geom = np.asarray(mesh.block_geom)
        dx, dy = float(geom[0, 4]), float(geom[0, 5])
        self.n, self.nfaces, self.dx, self.dy = n, nf, dx, dy
        self.volume = dx * dy
        self.cells = (xminus, xplus, yminus, yplus)
        right = fp[xplus]
        above = fp[yplus]
        self.comp = []
        for a in (0, 1):
            g2l = np.full(nf, -1)
            l2g = np.flatnonzero(ax == a)
            g2l[l2g] = np.arange(len(l2g))
            if a == 0:
                # normal dual faces through cell centres, then transverse ones
                # through each cell's upper-right vertex
                minus = np.concatenate([g2l[xminus], g2l[xplus]])
                plus = np.concatenate([g2l[xplus], g2l[xplus[above]]])
                fa = np.concatenate([xminus, yplus])
                fb = np.concatenate([xplus, yplus[right]])
                dax = np.concatenate([np.full(n, 1), np.full(n, 2)])
                area = np.concatenate([np.full(n, dy), np.full(n, dx)])
                dist = np.concatenate([np.full(n, dx), np.full(n, dy)])
            else:
                minus = np.concatenate([g2l[yminus], g2l[yplus]])
                plus = np.concatenate([g2l[yplus], g2l[yplus[right]]])
                fa = np.concatenate([yminus, xplus])
                fb = np.concatenate([yplus, xplus[above]])
                dax = np.concatenate([np.full(n, 2), np.full(n, 1)])
                area = np.concatenate([np.full(n, dx), np.full(n, dy)])
                dist = np.concatenate([np.full(n, dy), np.full(n, dx)])
            if np.any(minus < 0) or np.any(plus < 0):
                raise ValueError("invalid directional face mapping")
            nout = len(l2g)
            # grouped output schedule: outputs in order of their appearance
            # as (minus, plus) of dual faces 1, 2, ...
            targets = np.empty(2 * len(minus), dtype=np.int64)
            targets[0::2] = minus
            targets[1::2] = plus
            faces = np.repeat(np.arange(len(minus)), 2)
            signs = np.tile(np.array([1.0, -1.0]), len(minus))
            order = np.argsort(targets, kind="stable")
            counts = np.bincount(targets, minlength=nout)
            offsets = np.concatenate([[0], np.cumsum(counts)])
            self.comp.append(dict(l2g=l2g, g2l=g2l, minus=minus, plus=plus, fa=fa, fb=fb, axis=dax, area=area,
                                  dist=dist, offsets=offsets, faces=faces[order], signs=signs[order]))

    def _gather(self, c, dual, out_scale):
        """out[o] = (sum over o's dual faces, grouped order, of sign*dual) * out_scale."""
        vals = c["signs"] * dual[c["faces"]]
        nout = len(c["l2g"])
        out = np.zeros(nout)
        off = c["offsets"]
        width = off[1:] - off[:-1]
        # every output has the same number of dual faces on a uniform mesh;
        # sum left to right from zero as the reference's loop does
        w = int(width[0])
        if np.any(width != w):
            raise AssertionError("non-uniform dual adjacency")
        v = vals.reshape(nout, w)
        acc = np.zeros(nout)
        for k in range(w):
            acc = acc + v[:, k]
        out[:] = acc * out_scale
        return out

    def directional_divergence(self, a, u, sweep):
        """``_directional_flux_divergence!`` (uniform): the sweep's dual flux
        divergence of the advecting velocity per unknown of component a."""
        c = self.comp[a]
        active = (sweep == 0) | (c["axis"] == sweep)
        flux = np.where(active, (u[c["fa"]] + u[c["fb"]]) * c["area"] / 2, 0.0)
        return self._gather(c, flux, 1.0 / self.volume)

    def convection(self, a, q, u, dt, scheme="lw", sweep=0):
        """``_convection!`` of component a: the conservative tendency of the
        component values q (indexed by component unknown) transported by the
        face velocity u (all interior faces)."""
        if scheme not in SCHEMES:
            raise ValueError(f"unsupported staggered advection scheme: {scheme}")
        c = self.comp[a]
        split = scheme == "lw" and dt != 0
        div = self.directional_divergence(a, u, sweep) if split else np.zeros(len(c["l2g"]))
        active = (sweep == 0) | (c["axis"] == sweep)
        flux = (u[c["fa"]] + u[c["fb"]]) * c["area"] / 2
        qm, qp = q[c["minus"]], q[c["plus"]]
        if scheme == "upwind":
            state = np.where(flux >= 0, qm, qp)
        elif scheme == "centred":
            state = (qm + qp) / 2
        else:
            speed = flux / c["area"]
            courant = np.abs(dt * speed / c["dist"])
            pos = flux >= 0
            up = np.where(pos, qm, qp)
            down = np.where(pos, qp, qm)
            state = ((1 - courant) * down + (1 + courant) * up) / 2 - \
                dt * (qm * div[c["minus"]] + qp * div[c["plus"]]) / 4
        dual = np.where(active, flux * state, 0.0)
        return self._gather(c, dual, 1.0 / self.volume)

    def rates(self, a, u):
        """``_accumulate_staggered_rates!`` (uniform): sum of |dual flux| / V."""
        c = self.comp[a]
        flux = np.abs((u[c["fa"]] + u[c["fb"]]) * c["area"] / 2)
        nout = len(c["l2g"])
        v = flux[c["faces"]].reshape(nout, -1)
        acc = np.zeros(nout)
        for k in range(v.shape[1]):
            acc = acc + v[:, k]
        return acc * (1.0 / self.volume)


class StaggeredAdvection:
    """``StaggeredVelocityAdvectionWorkspace``: the schedule plus the sweep
    order flag that alternates x/y and y/x on successive ``advect!`` calls."""

    def __init__(self, mesh):
        self.schedule = UniformSchedule(mesh)
        self.sweep_forward = True

    def advect(self, field, advecting, dt, scheme="lw"):
        """``advect!(field, advecting, dt, workspace)`` on interior face
        values (both components); returns the new field. ``advecting is
        field`` gives the one-velocity shorthand (each published intermediate
        transports the next sweep)."""
        S = self.schedule
        if dt < 0:
            raise ValueError("dt must be nonnegative")
        field = np.array(field, np.float64)
        self_adv = advecting is None
        adv = field if self_adv else np.asarray(advecting, np.float64)
        wx = field[S.comp[0]["l2g"]].copy()
        wy = field[S.comp[1]["l2g"]].copy()
        split = scheme == "lw"
        stages = ((1, 2) if self.sweep_forward else (2, 1)) if split else (0,)
        for sweep in stages:
            cur = field if self_adv else adv
            tx = S.convection(0, wx, cur, dt, scheme, sweep)
            ty = S.convection(1, wy, cur, dt, scheme, sweep)
            wx = wx - dt * tx
            wy = wy - dt * ty
            if self_adv:
                field = field.copy()
                field[S.comp[0]["l2g"]] = wx
                field[S.comp[1]["l2g"]] = wy
        field[S.comp[0]["l2g"]] = wx
        field[S.comp[1]["l2g"]] = wy
        if split and dt != 0:
            self.sweep_forward = not self.sweep_forward
        return field

    def stable_timestep(self, u, cfl=0.9):
        if not (0 < cfl <= 1):
            raise ValueError("cfl must lie in (0, 1]")
        S = self.schedule
        out = np.inf
        for a in (0, 1):
            m = S.rates(a, np.asarray(u, np.float64)).max(initial=0.0)
            out = min(out, np.inf if m == 0 else cfl / m)
        return out


class StaggeredDiffusion:
    """``StaggeredDiffusionOperator(mesh; viscosity)`` on a uniform periodic
    mesh: per component the stiffness K (sum over dual connections of
    measure*distance * g g^T with the two-point derivative g = [-1/d, 1/d],
    symmetrised), the dual volumes V and the Laplacian L = -V^-1 K."""

    def __init__(self, mesh, viscosity):
        if not viscosity > 0:
            raise ValueError("constant viscosity must be strictly positive")
        self.schedule = S = UniformSchedule(mesh)
        self.viscosity = float(viscosity)
        self.K, self.L, self.V = [], [], []
        for a in (0, 1):
            c = S.comp[a]
            n = len(c["l2g"])
            inv_d = 1.0 / c["dist"]
            weight = c["area"] * c["dist"]
            st = np.stack([-inv_d, inv_d], axis=1)
            idx = np.stack([c["minus"], c["plus"]], axis=1)
            rows, cols, vals = [], [], []
            for lc in range(2):
                for lr in range(2):
                    rows.append(idx[:, lr])
                    cols.append(idx[:, lc])
                    vals.append(weight * st[:, lr] * st[:, lc])
            K = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                              shape=(n, n)).tocsc()
            K = sp.csc_matrix((K + K.T) / 2)
            K.eliminate_zeros()
            V = np.full(n, S.volume)
            self.K.append(K)
            self.V.append(V)
            self.L.append(sp.csc_matrix(-sp.diags(1.0 / V) @ K))

    def apply(self, u):
        """``diffusion!``: viscosity * L u per component, on interior faces."""
        u = np.asarray(u, np.float64)
        out = np.zeros_like(u)
        for a in (0, 1):
            l2g = self.schedule.comp[a]["l2g"]
            out[l2g] = self.viscosity * (self.L[a] @ u[l2g])
        return out

    def shifted(self, a, diffusion, reaction=1.0):
        """``_staggered_stable_matrix``: diffusion * K + reaction * V."""
        return sp.csc_matrix(diffusion * self.K[a] + sp.diags(reaction * self.V[a]))

    def implicit_midpoint(self, convected, old, dt):
        """``implicit_midpoint_diffusion!``: (I - dt nu L / 2) u = convected +
        dt nu L old / 2, i.e. per component (V + (dt nu / 2) K) u = V w."""
        if not dt > 0:
            raise ValueError("dt must be strictly positive")
        half = dt * self.viscosity / 2
        work = self.apply(old)
        w = np.asarray(convected, np.float64) + dt * work / 2
        out = np.zeros_like(w)
        for a in (0, 1):
            l2g = self.schedule.comp[a]["l2g"]
            M = self.shifted(a, half)
            out[l2g] = spla.spsolve(M, self.V[a] * w[l2g])
        return out


def step(u, p, solver, advection, diffusion, dt, *, force=None, scheme="lw", tolerance=1e-12, max_cycles=100):
    """``step!`` (src/single_phase/stepping.jl:194-264) of one phase on a
    uniform periodic mesh: transport with the old velocity frozen, the body
    force sampled at the midpoint time (face values ``force``), the
    implicit-midpoint viscous substep, the pressure predictor with the old
    pressure and the incremental projection. ``solver`` is the oracle's
    pressure solver (``OSolver`` with face coefficient 1/rho). Returns
    (u, p, projection stats)."""
    u = np.asarray(u, np.float64)
    old = u.copy()
    bound = advection.stable_timestep(old, cfl=1.0)
    if dt > bound * (1 + 1.0e-6 + 32 * np.finfo(float).eps):
        raise ValueError(f"time step {dt} exceeds advective stability bound {bound}")
    convected = advection.advect(old.copy(), old, dt, scheme)
    if force is not None:
        convected = convected + dt * np.asarray(force, np.float64)
    un = diffusion.implicit_midpoint(convected, old, dt)
    un = un - dt * solver.weighted_gradient(np.asarray(p, np.float64))
    un, pn, phi, st = solver.project(un, p, dt, tolerance=tolerance, max_cycles=max_cycles)
    return un, pn, st


def lax_wendroff_1d(q, c):
    """Classical one-dimensional Lax-Wendroff update of a periodic line with
    Courant number c (closed form used to pin the staggered sweep)."""
    qp = np.roll(q, -1)
    qm = np.roll(q, 1)
    return q - c / 2 * (qp - qm) + c * c / 2 * (qp - 2 * q + qm)
