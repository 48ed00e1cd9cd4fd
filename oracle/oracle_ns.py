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
  force ``_add_body_force!`` (src/navier_stokes/stages.jl:40-53);
- viscous preconditioner ``TreeStaggeredGeometricMultigridPreconditioner``
  (src/navier_stokes/implicit_diffusion.jl:299-541) on uniform periodic
  meshes: the staggered tree-coarsening plan (leaf pruning, then 2x2 root
  merges, src/solvers/multigrid/tree_coarsening.jl:7-201), rediscretised
  shifted operators (fine level ``V (I - h L)``, coarse ``h K + V``,
  implicit_diffusion.jl:213-261), the face prolongation restricted to each
  component (``_tree_face_prolongation``, tree_coarsening.jl:462-530:
  coarse-line faces take the coarse face, mid-line faces the mean of the two
  bracketing coarse faces), damped Jacobi with ``_jacobi_weight!``
  (hierarchy.jl:8-30, 786-817), the V-cycle ``_vcycle!``
  (cycles.jl:1-77, 262-283) and the dense coarse LU;
- the viscous Krylov solve ``_solve_krylov_with_stats!`` with
  ``relative_to_rhs`` around Krylov.jl's preconditioned ``cg``
  (src/solvers/multigrid/elliptic_solver.jl:462-627, cycles.jl:78-101).

``StaggeredDiffusion.implicit_midpoint`` solves the viscous system exactly
(sparse LU); ``implicit_midpoint_mg`` runs the reference's preconditioned
Krylov solve and reports its statistics.
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
        # lattice position of every cell, and the cell at every lattice
        # position (row-major)
        cx, cy = mesh.cell_centers()
        ox, oy = float(np.min(geom[:, 0])), float(np.min(geom[:, 1]))
        ci = np.rint((cx - ox) / dx - 0.5).astype(np.int64)
        cj = np.rint((cy - oy) / dy - 0.5).astype(np.int64)
        nlx = int(ci.max()) + 1
        self.cell_at = np.empty(n, np.int64)
        self.cell_at[cj * nlx + ci] = np.arange(n)
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


def staggered_levels(base_mesh, block_size, extent, level=0, coarse_size=64, max_levels=32):
    """Meshes of the staggered hierarchy of a uniform periodic mesh (all
    leaves at `level`): ``_tree_next_coarse_mesh!`` with
    ``_TreeCoarseningPlan{:staggered}``: prune leaf levels, then halve the
    base mesh while the root layout is even and every periodic axis keeps at
    least four base cells; stop when the unknowns (cells, at the root level)
    reach `coarse_size`, at `max_levels`, or without progress. Returns
    (base_mesh, level) per hierarchy level."""
    out = [(tuple(base_mesh), level)]
    cur_base, cur_level = tuple(base_mesh), level
    cells = lambda b, l: b[0] * b[1] * 4 ** l  # noqa: E731
    while cells(cur_base, cur_level) > coarse_size and len(out) < max_levels:
        if cur_level == 0:
            layout = (cur_base[0] // block_size[0], cur_base[1] // block_size[1])
            if layout[0] % 2 or layout[1] % 2:
                break
            nb = (cur_base[0] // 2, cur_base[1] // 2)
            if nb[0] < 4 or nb[1] < 4:
                break
            cur_base = nb
        else:
            cur_level -= 1
        out.append((cur_base, cur_level))
    return out


def face_prolongation(coarse_n, a):
    """Component-`a` face prolongation between nested uniform periodic grids
    (fine n = 2 coarse n per axis), unknowns row-major (j, i): along the
    component's normal axis even fine lines take the coarse face of the same
    line, odd ones the mean of the two bracketing coarse lines; along the
    tangential axis each coarse face covers two fine faces."""
    ncx, ncy = coarse_n
    nfx, nfy = 2 * ncx, 2 * ncy
    rows, cols, vals = [], [], []
    for jf in range(nfy):
        for if_ in range(nfx):
            r = jf * nfx + if_
            if a == 0:
                jc = jf // 2
                if if_ % 2 == 0:
                    rows.append(r); cols.append(jc * ncx + if_ // 2); vals.append(1.0)  # noqa: E702
                else:
                    lo = if_ // 2
                    rows += [r, r]; cols += [jc * ncx + lo, jc * ncx + (lo + 1) % ncx]; vals += [0.5, 0.5]  # noqa: E702
            else:
                ic = if_ // 2
                if jf % 2 == 0:
                    rows.append(r); cols.append((jf // 2) * ncx + ic); vals.append(1.0)  # noqa: E702
                else:
                    lo = jf // 2
                    rows += [r, r]; cols += [lo * ncx + ic, ((lo + 1) % ncy) * ncx + ic]; vals += [0.5, 0.5]  # noqa: E702
    return sp.csc_matrix((vals, (rows, cols)), shape=(nfx * nfy, ncx * ncy))


class StaggeredMultigrid:
    """TreeStaggeredGeometricMultigridPreconditioner of one component on a
    uniform periodic mesh, for the shifted system (V + h K) u = V w with
    h = dt nu / 2. Unknowns are the component's faces in row-major (j, i)
    lattice order (the x-face at the left, the y-face below cell (i, j))."""

    DENSE = 256

    def __init__(self, n, spacing, viscosity, half, levels, pre_sweeps=2, post_sweeps=2):
        self.pre, self.post = pre_sweeps, post_sweeps
        self.half = half
        self.levels = []
        nx, ny = n
        hx, hy = spacing
        for k in range(levels):
            Kx = ((hy * hx) * (1.0 / hx)) * (1.0 / hx)
            Ky = ((hx * hy) * (1.0 / hy)) * (1.0 / hy)
            V = hx * hy
            N = nx * ny
            idx = np.arange(N).reshape(ny, nx)
            rows, cols, kv = [], [], []
            for (di, dj, c) in ((1, 0, Kx), (0, 1, Ky)):
                a = idx.ravel()
                b = np.roll(np.roll(idx, -dj, axis=0), -di, axis=1).ravel()
                for (r_, c_, v_) in ((a, a, c), (b, b, c), (a, b, -c), (b, a, -c)):
                    rows.append(r_); cols.append(c_); kv.append(np.full(N, v_))  # noqa: E702
            K = sp.coo_matrix((np.concatenate(kv), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N)).tocsc()
            K = sp.csc_matrix((K + K.T) / 2)
            Vv = np.full(N, V)
            if k == 0:
                # the Krylov operator: volumes .* (reaction I - diffusion L)
                L = sp.csc_matrix(-sp.diags(1.0 / Vv) @ K)
                M = sp.csc_matrix(sp.diags(Vv) @ (sp.identity(N, format="csc") - half * L))
            else:
                M = sp.csc_matrix(half * K + sp.diags(Vv))
            d = M.diagonal()
            rowsum = np.asarray(abs(M).sum(axis=1)).ravel()
            bound = np.max(rowsum / d)
            w = min(1.0, 4.0 / (3.0 * bound))
            self.levels.append(dict(n=(nx, ny), M=M, dinv=1.0 / d, w=w, V=Vv))
            nx, ny, hx, hy = nx // 2, ny // 2, 2 * hx, 2 * hy
        for k in range(levels - 1):
            self.levels[k]["P"] = None  # set per component by `component`
        coarse = self.levels[-1]["M"]
        self.coarse_lu = spla.splu(sp.csc_matrix(coarse))

    def component(self, a):
        for k in range(len(self.levels) - 1):
            self.levels[k]["P"] = face_prolongation(self.levels[k + 1]["n"], a)
        return self

    def vcycle(self, b, k=0):
        L = self.levels[k]
        if k == len(self.levels) - 1:
            return self.coarse_lu.solve(b)
        x = np.zeros_like(b)
        for _ in range(self.pre):
            x = x + L["w"] * L["dinv"] * (b - L["M"] @ x)
        r = b - L["M"] @ x
        xc = self.vcycle(L["P"].T @ r, k + 1)
        x = x + L["P"] @ xc
        for _ in range(self.post):
            x = x + L["w"] * L["dinv"] * (b - L["M"] @ x)
        return x


def krylov_solve(M, V, b, x, precond, *, tol, max_cycles):
    """``_solve_krylov_with_stats!`` with ``relative_to_rhs`` around Krylov.jl's
    preconditioned CG (strong-residual callback, restart-and-verify loop,
    guarded round-off floor). Returns (x, stats)."""
    eps = np.finfo(float).eps

    def floor_of(x):
        return 128 * eps * max(1.0, np.max((abs(M) @ np.abs(x)) / V), np.max(np.abs(b / V)))

    def guarded(target, initial, current, x):
        if not (np.isfinite(initial) and np.isfinite(current) and target < current < initial):
            return target
        f = floor_of(x)
        return max(target, f) if (np.isfinite(f) and f < initial / 100) else target

    r = b - M @ x
    initial = np.max(np.abs(r / V))
    target = max(tol * np.max(np.abs(b / V)), floor_of(x))
    its = 0
    if initial > target:
        while its < max_cycles:
            r = b - M @ x
            z = precond(r)
            p = z.copy()
            gamma = r @ z
            run = 0
            if gamma != 0:
                while its + run < max_cycles:
                    Ap = M @ p
                    alpha = gamma / (p @ Ap)
                    x = x + alpha * p
                    r = r - alpha * Ap
                    run += 1
                    if np.max(np.abs(r / V)) <= target:
                        break
                    z = precond(r)
                    g2 = r @ z
                    p = z + (g2 / gamma) * p
                    gamma = g2
            its += run
            cur = np.max(np.abs((b - M @ x) / V))
            if cur <= guarded(target, initial, cur, x) or run == 0 or not np.isfinite(cur):
                break
    final = np.max(np.abs((b - M @ x) / V))
    t = guarded(target, initial, final, x)
    return x, dict(cycles=its, converged=bool(final <= t), initial_linf=initial, final_linf=final, target_linf=t)


def implicit_midpoint_mg(diffusion, mesh_params, convected, old, dst, dt, *, tol, max_cycles=100, coarse_size=64):
    """``implicit_midpoint_diffusion!`` with the reference's solver: per
    component the Krylov solve of V (I - h L) u = V w preconditioned by the
    staggered tree multigrid, from the initial guess `dst`. `mesh_params`:
    (base_mesh, block_size, level) of the uniform periodic mesh. Returns
    (u, [stats_x, stats_y])."""
    S = diffusion.schedule
    half = dt * diffusion.viscosity / 2
    base, block, level = mesh_params
    nlev = len(staggered_levels(base, block, None, level, coarse_size))
    nx, ny = base[0] * 2 ** level, base[1] * 2 ** level
    mg = StaggeredMultigrid((nx, ny), (S.dx, S.dy), diffusion.viscosity, half, nlev)
    work = diffusion.apply(old)
    w = np.asarray(convected, np.float64) + dt * work / 2
    out = np.array(dst, np.float64)
    stats = []
    for a in (0, 1):
        perm = lattice_order(diffusion, a, (nx, ny))
        l2g = S.comp[a]["l2g"][perm]
        comp = mg.component(a)
        M = comp.levels[0]["M"]
        V = comp.levels[0]["V"]
        x, st = krylov_solve(M, V, V * w[l2g], out[l2g].copy(), comp.vcycle, tol=tol, max_cycles=max_cycles)
        out[l2g] = x
        stats.append(st)
    return out, stats


def lattice_order(diffusion, a, n):
    """Permutation taking component `a`'s unknowns (mesh face order) to the
    row-major lattice order of its faces: the unknown of lattice cell (i, j)
    is that cell's minus face."""
    S = diffusion.schedule
    xminus, _, yminus, _ = S.cells
    faces = xminus if a == 0 else yminus
    return S.comp[a]["g2l"][faces[S.cell_at]]


def step(u, p, solver, advection, diffusion, dt, *, force=None, scheme="lw", tolerance=1e-12, max_cycles=100,
         mg_params=None, coarse_size=64, viscous_stats=None):
    """``step!`` (src/single_phase/stepping.jl:194-264) of one phase on a
    uniform periodic mesh: transport with the old velocity frozen, the body
    force sampled at the midpoint time (face values ``force``), the
    implicit-midpoint viscous substep, the pressure predictor with the old
    pressure and the incremental projection. ``solver`` is the oracle's
    pressure solver (``OSolver`` with face coefficient 1/rho). The viscous
    substep is exact, or with ``mg_params`` = (base_mesh, block_size, level)
    the reference's multigrid-preconditioned Krylov solve (its statistics
    appended to ``viscous_stats``). Returns (u, p, projection stats)."""
    u = np.asarray(u, np.float64)
    old = u.copy()
    bound = advection.stable_timestep(old, cfl=1.0)
    if dt > bound * (1 + 1.0e-6 + 32 * np.finfo(float).eps):
        raise ValueError(f"time step {dt} exceeds advective stability bound {bound}")
    convected = advection.advect(old.copy(), old, dt, scheme)
    if force is not None:
        convected = convected + dt * np.asarray(force, np.float64)
    if mg_params is None:
        un = diffusion.implicit_midpoint(convected, old, dt)
    else:
        un, vst = implicit_midpoint_mg(diffusion, mg_params, convected, old, old, dt, tol=tolerance,
                                       max_cycles=max_cycles, coarse_size=coarse_size)
        if not all(v["converged"] for v in vst):
            raise ValueError("staggered diffusion solve did not converge")
        if viscous_stats is not None:
            viscous_stats.append(vst)
    un = un - dt * solver.weighted_gradient(np.asarray(p, np.float64))
    un, pn, phi, st = solver.project(un, p, dt, tolerance=tolerance, max_cycles=max_cycles)
    return un, pn, st


def lax_wendroff_1d(q, c):
    """Classical one-dimensional Lax-Wendroff update of a periodic line with
    Courant number c (closed form used to pin the staggered sweep)."""
    qp = np.roll(q, -1)
    qm = np.roll(q, 1)
    return q - c / 2 * (qp - qm) + c * c / 2 * (qp - 2 * q + qm)
