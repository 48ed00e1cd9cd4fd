"""The oracle's one-velocity momentum step (oracle/oracle_ns.py) pinned by the
reference's staggered-advection and Navier-Stokes tests and by closed forms:
classical one-dimensional Lax-Wendroff lines (with the sweep order of two
successive calls), conservation, the skew-symmetric centred operator,
independence of the block partition, the discrete Laplacian's Fourier
eigenvalues, the implicit-midpoint amplification factor, the CFL bound and
the Taylor-Green convergence rate of whole steps."""
import numpy as np
import pytest

from oracle import OMesh, OSolver
from oracle_ns import StaggeredAdvection, StaggeredDiffusion, UniformSchedule, lax_wendroff_1d, step


def face_grid(mesh, a):
    """Component a's unknowns as (x index, y index) of a uniform grid, from
    the face centres."""
    sch = UniformSchedule(mesh)
    l2g = sch.comp[a]["l2g"]
    c = mesh.face_center[l2g]
    ix = np.rint((c[:, 0] - (0.0 if a == 0 else sch.dx / 2)) / sch.dx).astype(int)
    iy = np.rint((c[:, 1] - (sch.dy / 2 if a == 0 else 0.0)) / sch.dy).astype(int)
    nx = ix.max() + 1
    ny = iy.max() + 1
    return l2g, ix % nx, iy % ny, nx, ny


def sample(mesh, fx, fy):
    u = np.zeros(mesh.nfaces)
    c = mesh.face_center
    ax = mesh.face_axis
    u[ax == 0] = fx(c[ax == 0, 0], c[ax == 0, 1])
    u[ax == 1] = fy(c[ax == 1, 0], c[ax == 1, 1])
    return u


def test_lax_wendroff_lines_and_sweep_order():
    """With u = a(y), v = b(x) every sweep moves each grid line with its own
    constant flux, so one advect! is classical 1D Lax-Wendroff along x then
    along y, and the next call applies the two sweeps in the other order
    (advect!, staggered_update.jl:265-326; 'On a uniform constant-velocity
    line it reduces to classical one-dimensional LW')."""
    m = OMesh(base_mesh=(16, 16), block_size=(8, 8), extent=(1.0, 1.0))
    ax_fun = lambda x, y: 0.7 + 0.3 * np.cos(2 * np.pi * y)  # noqa: E731
    by_fun = lambda x, y: -0.4 + 0.2 * np.sin(2 * np.pi * x)  # noqa: E731
    u = sample(m, ax_fun, by_fun)
    rng = np.random.default_rng(3)
    q = rng.standard_normal(m.nfaces)
    dt = 0.01
    adv = StaggeredAdvection(m)
    out1 = adv.advect(q.copy(), u, dt)
    out2 = adv.advect(out1.copy(), u, dt)
    h = 1.0 / 16
    for a in (0, 1):
        l2g, ix, iy, nx, ny = face_grid(m, a)
        G = np.zeros((ny, nx))
        G[iy, ix] = q[l2g]
        # line Courant numbers: the flux through each dual face of a line
        xs = (np.arange(nx) + 0.5) * h  # cell centres
        ys = (np.arange(ny) + 0.5) * h
        if a == 0:
            # x-sweep: normal dual faces, flux (u + u) dy / 2 with u = a(y_j)
            cx = dt * ax_fun(0, ys) / h
            # y-sweep: transverse dual faces at x = i h average v on the two
            # y-faces left and right of the vertex: b(x_{i-1/2}), b(x_{i+1/2})
            cy = dt * (by_fun(np.roll(xs, 1), 0) + by_fun(xs, 0)) / 2 / h
        else:
            # x-sweep: transverse dual faces at y = j h average u on the x-faces
            # below and above the vertex
            cx = dt * (ax_fun(0, np.roll(ys, 1)) + ax_fun(0, ys)) / 2 / h
            cy = dt * by_fun(xs, 0) / h

        def sweep_x(F):
            return np.stack([lax_wendroff_1d(F[j], cx[j]) for j in range(ny)])

        def sweep_y(F):
            return np.stack([lax_wendroff_1d(F[:, i], cy[i]) for i in range(nx)], axis=1)

        E1 = sweep_y(sweep_x(G))
        E2 = sweep_x(sweep_y(E1))
        got1 = np.zeros((ny, nx))
        got1[iy, ix] = out1[l2g]
        got2 = np.zeros((ny, nx))
        got2[iy, ix] = out2[l2g]
        assert np.max(np.abs(got1 - E1)) < 1e-13
        assert np.max(np.abs(got2 - E2)) < 1e-13
        # the two orders differ on this field: the alternation is observable
        assert np.max(np.abs(sweep_x(sweep_y(G)) - E1)) > 1e-6


def test_transport_conserves_each_component():
    m = OMesh(base_mesh=(32, 32), block_size=(16, 16), extent=(2.0, 2.0))
    u = sample(m, lambda x, y: np.sin(np.pi * x) * np.cos(np.pi * y), lambda x, y: -np.cos(np.pi * x) * np.sin(np.pi * y))
    q = np.cos(np.arange(m.nfaces) * 0.37) + 2.0
    adv = StaggeredAdvection(m)
    out = adv.advect(q.copy(), u, 0.01)
    for a in (0, 1):
        l2g = adv.schedule.comp[a]["l2g"]
        assert abs(out[l2g].sum() - q[l2g].sum()) < 1e-12 * np.abs(q[l2g]).sum()


def test_centred_convection_matrix_is_skew_symmetric():
    """'centred convection matrix' (test/operators/advection/staggered_advection.jl:1298-1319)."""
    m = OMesh(block_size=(4, 4))
    u = sample(m, lambda x, y: np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y),
               lambda x, y: -np.cos(2 * np.pi * x) * np.sin(2 * np.pi * y))
    S = UniformSchedule(m)
    for a in (0, 1):
        n = len(S.comp[a]["l2g"])
        A = np.zeros((n, n))
        for s in range(n):
            e = np.zeros(n)
            e[s] = 1
            A[:, s] = S.convection(a, e, u, 0.0, "centred")
        assert np.allclose(A, -A.T, atol=2e-14, rtol=2e-14)
        assert np.max(np.abs(S.convection(a, np.ones(n), u, 0.0, "centred"))) <= 64 * np.finfo(float).eps


def test_lax_wendroff_translation_converges_at_second_order():
    """'Lax-Wendroff convergence' (staggered_advection.jl:1321-1351)."""
    def err(nx):
        m = OMesh(block_size=(nx, 4))
        speed = 0.5
        u = sample(m, lambda x, y: np.full_like(x, speed), lambda x, y: np.zeros_like(x))
        q = sample(m, lambda x, y: np.sin(2 * np.pi * x), lambda x, y: np.zeros_like(x))
        adv = StaggeredAdvection(m)
        final = 0.2
        steps = int(np.ceil(final / (0.4 / (nx * speed))))
        dt = final / steps
        for _ in range(steps):
            q = adv.advect(q, u, dt)
        l2g = adv.schedule.comp[0]["l2g"]
        exact = np.sin(2 * np.pi * (m.face_center[l2g, 0] - speed * final))
        return np.sqrt(np.mean((q[l2g] - exact) ** 2))

    e = [err(n) for n in (16, 32, 64)]
    assert np.log2(e[0] / e[1]) > 1.9
    assert np.log2(e[1] / e[2]) > 1.9


def by_coordinate(m, v):
    keys = np.round(np.column_stack([m.face_axis, m.face_center]), 12)
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
    return v[order]


def test_transport_is_independent_of_block_partition():
    """'staggered transport is independent of uniform block partitioning'
    (staggered_advection.jl:1388-1415): bitwise equal results."""
    meshes = [OMesh(block_size=(8, 8)),
              OMesh(block_size=(4, 4), max_level=1, refine=lambda k, o, e: True),
              OMesh(base_mesh=(8, 8), block_size=(4, 4))]
    res = []
    for m in meshes:
        u = sample(m, lambda x, y: np.full_like(x, 0.3), lambda x, y: np.full_like(x, -0.15))
        q = sample(m, lambda x, y: 1 + 0.2 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y),
                   lambda x, y: 1 + 0.2 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y))
        res.append(by_coordinate(m, StaggeredAdvection(m).advect(q, u, 0.02)))
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])


def test_viscous_operator_fourier_eigenvalues_and_midpoint_factor():
    """The staggered Laplacian of a periodic Fourier mode is -(4/h^2)
    sin^2(k h / 2) times the mode (per direction); the implicit-midpoint
    substep multiplies it by (1 + dt nu lam / 2) / (1 - dt nu lam / 2)."""
    m = OMesh(base_mesh=(32, 32), block_size=(16, 16), extent=(2 * np.pi, 2 * np.pi))
    nu, dt = 0.02, 1e-3
    D = StaggeredDiffusion(m, nu)
    h = 2 * np.pi / 32
    kx, ky = 3, 2
    u = sample(m, lambda x, y: np.sin(kx * x) * np.cos(ky * y), lambda x, y: np.cos(kx * x) * np.sin(ky * y))
    lam = -(4 / h ** 2) * (np.sin(kx * h / 2) ** 2 + np.sin(ky * h / 2) ** 2)
    Lu = D.apply(u)
    assert np.max(np.abs(Lu - nu * lam * u)) < 1e-11
    out = D.implicit_midpoint(u, u, dt)
    g = (1 + dt * nu * lam / 2) / (1 - dt * nu * lam / 2)
    assert np.max(np.abs(out - g * u)) < 1e-13
    # stiffness rows sum to zero and the matrix is symmetric
    for a in (0, 1):
        K = D.K[a]
        assert np.max(np.abs(K.sum(axis=1))) < 1e-12
        assert abs(K - K.T).max() == 0.0


def test_stable_timestep_of_a_constant_velocity():
    """Each unknown's rate is the sum of |dual flux| over its four dual faces
    divided by its dual volume: 2(|a| dy + |b| dx) / (dx dy)."""
    m = OMesh(base_mesh=(16, 16), block_size=(8, 8), extent=(2.0, 1.0))
    a, b = 0.6, -0.25
    u = sample(m, lambda x, y: np.full_like(x, a), lambda x, y: np.full_like(x, b))
    dx, dy = 2.0 / 16, 1.0 / 16
    rate = 2 * (abs(a) * dy + abs(b) * dx) / (dx * dy)
    assert abs(StaggeredAdvection(m).stable_timestep(u, cfl=0.9) - 0.9 / rate) < 1e-14


def taylor_green_error(nx):
    """``_taylor_green_error(nx)`` (test/navier_stokes/navier_stokes.jl:658-689):
    one nx x nx root block on the periodic [0, 2 pi]^2, nu = 0.02, nx steps of
    dt = 0.02 / nx, tolerance 2e-10; the dual-volume weighted L2 error of the
    velocity against the decaying exact field (:7-21) and the largest cell
    divergence."""
    m = OMesh(block_size=(nx, nx), extent=(2 * np.pi, 2 * np.pi))
    nu = 0.02
    exact = lambda t: (lambda x, y: np.exp(-2 * nu * t) * np.sin(x) * np.cos(y),  # noqa: E731
                       lambda x, y: -np.exp(-2 * nu * t) * np.cos(x) * np.sin(y))
    u = sample(m, *exact(0.0))
    p = np.zeros(m.ncells)
    S = OSolver(m, np.ones(m.nfaces), coarse_size=16)
    adv = StaggeredAdvection(m)
    D = StaggeredDiffusion(m, nu)
    dt = 0.02 / nx
    for _ in range(nx):
        u, p, st = step(u, p, S, adv, D, dt, tolerance=2e-10, max_cycles=80)
        assert st["converged"]
    ex = sample(m, *exact(0.02))
    V = adv.schedule.volume
    err = np.sqrt(np.sum(V * (u - ex) ** 2) / (V * m.nfaces))
    return err, np.max(np.abs(m.divergence(u)))


def test_taylor_green_steps_converge_at_second_order():
    """'Taylor--Green Navier--Stokes integration'
    (test/navier_stokes/navier_stokes.jl:691-704): the velocity error falls at
    second order from 12^2 to 24^2 cells and every step leaves a discretely
    divergence-free velocity."""
    coarse = taylor_green_error(12)
    fine = taylor_green_error(24)
    assert fine[0] < coarse[0]
    assert np.log2(coarse[0] / fine[0]) > 1.8
    assert coarse[1] <= 2e-8 and fine[1] <= 2e-8


def test_step_rejects_a_step_beyond_the_advective_bound():
    m = OMesh(base_mesh=(16, 16), block_size=(8, 8))
    u = sample(m, lambda x, y: np.full_like(x, 2.0), lambda x, y: np.zeros_like(x))
    adv = StaggeredAdvection(m)
    D = StaggeredDiffusion(m, 0.01)
    bound = adv.stable_timestep(u, cfl=1.0)
    with pytest.raises(ValueError, match="stability bound"):
        step(u, np.zeros(m.ncells), OSolver(m, np.ones(m.nfaces)), adv, D, 1.01 * bound)


def test_staggered_multigrid_hierarchy_and_transfer():
    """'staggered base-mesh multigrid' (test/operators/staggered_diffusion.jl:
    1752-1800), periodic case: base meshes (16,16), (8,8), (4,4) with
    coarse_size 16, P 1 = 1 and P' V_fine = V_coarse; a single root block has
    no level to prune ('implicit midpoint staggered elliptic solve', :1694-1750:
    one level); leaf levels are pruned before the base mesh is halved."""
    from oracle_ns import face_prolongation, staggered_levels
    assert staggered_levels((16, 16), (4, 4), None, 0, 16) == [((16, 16), 0), ((8, 8), 0), ((4, 4), 0)]
    assert staggered_levels((16, 16), (16, 16), None, 0, 64) == [((16, 16), 0)]
    assert staggered_levels((16, 16), (4, 4), None, 1, 64) == [((16, 16), 1), ((16, 16), 0), ((8, 8), 0)]
    assert staggered_levels((128, 128), (16, 16), None, 0, 64) == [((128, 128), 0), ((64, 64), 0), ((32, 32), 0),
                                                                   ((16, 16), 0)]
    for a in (0, 1):
        for nc in ((8, 8), (4, 4)):
            P = face_prolongation(nc, a)
            nf = 4 * nc[0] * nc[1]
            assert np.max(np.abs(P @ np.ones(nc[0] * nc[1]) - 1)) <= 64 * np.finfo(float).eps
            h = 1.0 / (2 * nc[0])
            assert np.max(np.abs(P.T @ np.full(nf, h * h) - 4 * h * h)) <= 64 * np.finfo(float).eps


def test_staggered_multigrid_solve():
    """The preconditioned viscous solve: a single 16^2 root (one level, the
    dense coarse LU) reproduces the reference's amplification test
    ('implicit midpoint staggered elliptic solve', staggered_diffusion.jl:
    1694-1728: a sin(2 pi x) / cos(2 pi y) field, viscosity 0.1, dt 0.1,
    2e-13); on a multi-level periodic mesh the V-cycle is symmetric and the
    solve equals the exact sparse solution."""
    from oracle_ns import StaggeredMultigrid, implicit_midpoint_mg
    m = OMesh(block_size=(16, 16))
    nu, dt = 0.1, 0.1
    D = StaggeredDiffusion(m, nu)
    old = sample(m, lambda x, y: np.sin(2 * np.pi * x), lambda x, y: np.cos(2 * np.pi * y))
    u, st = implicit_midpoint_mg(D, ((16, 16), (16, 16), 0), old, old, np.zeros(m.nfaces), dt,
                                 tol=np.sqrt(np.finfo(float).eps))
    dx = 1.0 / 16
    lam = -4 * np.sin(np.pi / 16) ** 2 / dx ** 2
    g = (1 + dt * nu * lam / 2) / (1 - dt * nu * lam / 2)
    assert np.max(np.abs(u - g * old)) <= 2e-13 * (1 + np.max(np.abs(old)))
    assert all(s["converged"] for s in st)
    m2 = OMesh(base_mesh=(32, 32), block_size=(8, 8), extent=(2 * np.pi, 2 * np.pi))
    D2 = StaggeredDiffusion(m2, 0.5)
    mg = StaggeredMultigrid((32, 32), (2 * np.pi / 32,) * 2, 0.5, 0.5 * 0.05 / 2, 3).component(0)
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal(1024), rng.standard_normal(1024)
    assert abs(x @ mg.vcycle(y) - mg.vcycle(x) @ y) <= 1e-10 * abs(x @ mg.vcycle(y))
    old2 = sample(m2, lambda x, y: np.sin(x) * np.cos(2 * y), lambda x, y: np.cos(3 * x))
    conv2 = old2 + 0.1 * np.cos(np.arange(m2.nfaces) * 0.3)
    u2, st2 = implicit_midpoint_mg(D2, ((32, 32), (8, 8), 0), conv2, old2, old2.copy(), 0.05, tol=1e-13)
    assert np.max(np.abs(u2 - D2.implicit_midpoint(conv2, old2, 0.05))) < 1e-12
    assert all(s["converged"] and s["cycles"] <= 8 for s in st2)
