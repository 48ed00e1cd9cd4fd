"""GPU parity of the one-velocity momentum step (staggered transport, viscous
operator, implicit-midpoint viscous solve, step!) against the oracle's
restatement on uniform periodic meshes, at the Taylor-Green configuration's
own size (128^2 cells, 16^2 blocks) among others.

Tolerances: the Lax-Wendroff transport is bit-exact (each unknown's sweep
tendency sums two dual fluxes, an order-free sum); the unsplit schemes and
the viscous operator sum five terms in another order than the oracle's
sparse product, within 1e-14 relative; the viscous solve runs to 1e-13
relative against the oracle's exact sparse solution; whole steps compare
the velocity to 1e-10 and the pressure to 1e-9 relative with both sides'
pressure solves at tolerance 1e-12."""
import numpy as np
import pytest

import tpamr
from oracle import OMesh, OSolver
from oracle_ns import StaggeredAdvection, StaggeredDiffusion, implicit_midpoint_mg, step as ostep

pytestmark = pytest.mark.gpu

MESHES = {
    "one_root_32": dict(base_mesh=(32, 32), block_size=(32, 32), extent=(1.0, 1.0)),
    "roots_8": dict(base_mesh=(32, 16), block_size=(8, 8), extent=(2.0, 1.0)),
    "refined_4": dict(base_mesh=(16, 16), block_size=(4, 4), extent=(1.0, 1.0), max_level=1,
                      refine=lambda k, o, e: True),
    "taylor_green_128": dict(base_mesh=(128, 128), block_size=(16, 16), extent=(2 * np.pi, 2 * np.pi)),
}


def pair(name):
    kw = dict(MESHES[name], periodic=(True, True))
    return OMesh(**kw), tpamr.BlockTreeMesh(**kw)


def sample(m, fx, fy):
    """Face values of (fx, fy) at the oracle mesh's face centres (the
    library's mesh has the same faces in the same order)."""
    u = np.zeros(m.nfaces)
    c = m.face_center
    ax = np.asarray(m.face_axis)
    u[ax == 0] = fx(c[ax == 0, 0], c[ax == 0, 1])
    u[ax == 1] = fy(c[ax == 1, 0], c[ax == 1, 1])
    return u


def tg(m, t=0.0, nu=0.02):
    e = np.exp(-2 * nu * t)
    return sample(m, lambda x, y: e * np.sin(x) * np.cos(y), lambda x, y: -e * np.cos(x) * np.sin(y))


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("name", list(MESHES))
def test_lax_wendroff_transport_is_bit_exact(name):
    om, pm = pair(name)
    rng = np.random.default_rng(11)
    Lx, Ly = MESHES[name]["extent"]
    adv_vel = sample(om, lambda x, y: 0.3 + 0.2 * np.sin(2 * np.pi * y / Ly),
                     lambda x, y: -0.2 + 0.1 * np.cos(2 * np.pi * x / Lx))
    q = rng.standard_normal(om.nfaces)
    dt = 0.25 * StaggeredAdvection(om).stable_timestep(adv_vel, cfl=1.0)
    sysg = tpamr.NavierStokesSystem(pm, dynamic_viscosity=0.01)
    oadv = StaggeredAdvection(om)
    got, want = q.copy(), q.copy()
    for call in range(3):
        assert sysg.sweep_forward == oadv.sweep_forward
        sysg.advect(got, adv_vel, dt)
        want = oadv.advect(want, adv_vel, dt)
        assert np.array_equal(got, want), (name, call, rel(got, want))
    # the one-velocity shorthand: each sweep transported by the published
    # intermediate
    u0 = 0.5 * adv_vel
    g1 = u0.copy()
    sysg.advect(g1, g1, dt)
    o1 = oadv.advect(u0.copy(), None, dt)
    assert np.array_equal(g1, o1)
    sysg.close()


@pytest.mark.parametrize("scheme", ["upwind", "centred"])
def test_unsplit_transport_matches_oracle(scheme):
    om, pm = pair("roots_8")
    u = sample(om, lambda x, y: np.sin(np.pi * x) * np.cos(2 * np.pi * y), lambda x, y: 0.3 * np.cos(np.pi * x))
    q = np.cos(np.arange(om.nfaces) * 0.17)
    dt = 0.3 * StaggeredAdvection(om).stable_timestep(u, cfl=1.0)
    sysg = tpamr.NavierStokesSystem(pm, dynamic_viscosity=0.01, scheme=scheme)
    got = q.copy()
    sysg.advect(got, u, dt)
    want = StaggeredAdvection(om).advect(q.copy(), u, dt, scheme)
    assert rel(got, want) < 1e-14
    assert sysg.sweep_forward  # unsplit schemes keep the order flag
    sysg.close()


@pytest.mark.parametrize("name", ["roots_8", "taylor_green_128"])
def test_viscous_operator_and_cfl_bound_match_oracle(name):
    om, pm = pair(name)
    nu = 0.02
    u = tg(om) + 0.1 * np.cos(np.arange(om.nfaces) * 0.31)
    sysg = tpamr.NavierStokesSystem(pm, density=2.0, dynamic_viscosity=2 * nu)
    D = StaggeredDiffusion(om, nu)
    assert rel(sysg.diffusion(u), D.apply(u)) < 1e-14
    for cfl in (0.45, 1.0):
        assert abs(sysg.stable_timestep(u, cfl=cfl) / StaggeredAdvection(om).stable_timestep(u, cfl=cfl) - 1) < 1e-14
    sysg.close()


@pytest.mark.parametrize("name", ["refined_4", "taylor_green_128"])
def test_implicit_midpoint_diffusion_matches_oracle(name):
    om, pm = pair(name)
    nu, dt = 0.05, 2e-3
    old = tg(om)
    conv = old + 0.01 * np.sin(np.arange(om.nfaces) * 0.7)
    sysg = tpamr.NavierStokesSystem(pm, dynamic_viscosity=nu)
    dst = old.copy()
    sx, sy = sysg.implicit_midpoint_diffusion(dst, conv, old, dt, tolerance=1e-13)
    want = StaggeredDiffusion(om, nu).implicit_midpoint(conv, old, dt)
    assert sx.converged and sy.converged and sx.cycles > 0
    assert sx.final_linf <= sx.target_linf and sy.final_linf <= sy.target_linf
    assert rel(dst, want) < 1e-12
    sysg.close()


def mg_params(name):
    kw = MESHES[name]
    return kw["base_mesh"], kw["block_size"], kw.get("max_level", 0)


def run_steps(om, pm, u0, steps, dt, nu, tol, force=None, name=None):
    sysg = tpamr.NavierStokesSystem(pm, dynamic_viscosity=nu)
    ug, pg = u0.copy(), np.zeros(pm.ncells)
    S = OSolver(om, np.ones(om.nfaces))
    oadv, D = StaggeredAdvection(om), StaggeredDiffusion(om, nu)
    uo, po = u0.copy(), np.zeros(om.ncells)
    stats = []
    for _ in range(steps):
        st = tpamr.step(ug, pg, sysg, dt, force=force, tolerance=tol)
        vo = []
        uo, po, so = ostep(uo, po, S, oadv, D, dt, force=force, tolerance=tol,
                           mg_params=None if name is None else mg_params(name), viscous_stats=vo)
        stats.append((st, so, sysg.viscous_stats(), vo[0] if vo else None))
    sysg.close()
    return ug, pg, uo, po, stats


def test_taylor_green_two_steps_field_for_field():
    """The plain-method FP64 configuration: Taylor-Green on the periodic
    [0, 2 pi]^2 with 128^2 cells in 16^2 blocks, nu = 0.02, dt = 1e-3, two
    steps, velocity and pressure against the oracle."""
    om, pm = pair("taylor_green_128")
    nu, dt = 0.02, 1e-3
    ug, pg, uo, po, stats = run_steps(om, pm, tg(om), 2, dt, nu, 1e-12, name="taylor_green_128")
    for st, so, (vx, vy), vo in stats:
        assert st.converged and so["converged"]
        assert st.cycles == so["cycles"]
        assert vx.converged and vy.converged
        # the viscous solves take the reference's multigrid iteration counts
        assert (vx.cycles, vy.cycles) == (vo[0]["cycles"], vo[1]["cycles"])
    assert rel(ug, uo) < 1e-10
    assert rel(pg, po) < 1e-9
    # the velocity stays discretely divergence-free and follows the decaying
    # exact field to the discretisation error
    assert np.max(np.abs(om.divergence(ug))) < 1e-8
    assert rel(ug, tg(om, 2 * dt, nu)) < 1e-3


def test_step_with_body_force_matches_oracle():
    om, pm = pair("roots_8")
    nu = 0.01
    u0 = sample(om, lambda x, y: np.sin(np.pi * x) * np.cos(2 * np.pi * y),
                lambda x, y: -0.5 * np.cos(np.pi * x) * np.sin(2 * np.pi * y))
    f = sample(om, lambda x, y: np.cos(2 * np.pi * y), lambda x, y: 0.3 * np.sin(np.pi * x))
    dt = 0.4 * StaggeredAdvection(om).stable_timestep(u0, cfl=1.0)
    ug, pg, uo, po, _ = run_steps(om, pm, u0, 3, dt, nu, 1e-12, force=f, name="roots_8")
    assert rel(ug, uo) < 1e-10 and rel(pg, po) < 1e-9


def tg_error_gpu(n):
    """The reference's Taylor-Green integration case at n^2 cells (one
    n x n root block, dt = 0.02 / n, n steps, tolerance 2e-10)."""
    kw = dict(base_mesh=(n, n), block_size=(n, n), extent=(2 * np.pi, 2 * np.pi), periodic=(True, True))
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    nu = 0.02
    sysg = tpamr.NavierStokesSystem(pm, dynamic_viscosity=nu, coarse_size=16)
    u, p = tg(om), np.zeros(pm.ncells)
    for _ in range(n):
        st = tpamr.step(u, p, sysg, 0.02 / n, tolerance=2e-10, max_cycles=80)
        assert st.converged
    sysg.close()
    ex = tg(om, 0.02, nu)
    return np.sqrt(np.mean((u - ex) ** 2)), np.max(np.abs(om.divergence(u)))


def test_taylor_green_integration_converges_at_second_order():
    """'Taylor--Green Navier--Stokes integration'
    (test/navier_stokes/navier_stokes.jl:691-704) on the library's block
    sizes (16^2 and 32^2 cells instead of 12^2 and 24^2)."""
    coarse, fine = tg_error_gpu(16), tg_error_gpu(32)
    assert fine[0] < coarse[0]
    assert np.log2(coarse[0] / fine[0]) > 1.8
    assert coarse[1] <= 2e-8 and fine[1] <= 2e-8


def test_step_device_equals_step():
    torch = pytest.importorskip("torch")
    om, pm = pair("roots_8")
    u0 = tg(om) * 0.3 + 0.05 * np.cos(np.arange(om.nfaces) * 0.13)
    p0 = np.cos(np.arange(pm.ncells) * 0.07)
    sh = tpamr.NavierStokesSystem(pm, dynamic_viscosity=0.01)
    sd = tpamr.NavierStokesSystem(pm, dynamic_viscosity=0.01)
    dt = 0.3 * sh.stable_timestep(u0, cfl=1.0)
    uh, ph = u0.copy(), p0.copy()
    ud = torch.tensor(u0, device="cuda:0")
    pd = torch.tensor(p0, device="cuda:0")
    for _ in range(2):
        a = tpamr.step(uh, ph, sh, dt, tolerance=1e-11)
        b = tpamr.step_device(ud, pd, sd, dt, tolerance=1e-11)
        assert a.cycles == b.cycles
    torch.cuda.synchronize()
    assert np.array_equal(uh, ud.cpu().numpy()) and np.array_equal(ph, pd.cpu().numpy())
    sh.close()
    sd.close()


def test_step_rejects_what_the_reference_rejects():
    om, pm = pair("roots_8")
    sysg = tpamr.NavierStokesSystem(pm, dynamic_viscosity=0.01)
    u = sample(om, lambda x, y: np.full_like(x, 2.0), lambda x, y: np.zeros_like(x))
    bound = sysg.stable_timestep(u, cfl=1.0)
    with pytest.raises(ValueError, match="stability bound"):
        tpamr.step(u.copy(), np.zeros(pm.ncells), sysg, 1.01 * bound)
    with pytest.raises(ValueError, match="cfl"):
        sysg.stable_timestep(u, cfl=1.5)
    sysg.close()
    walls = tpamr.BlockTreeMesh(base_mesh=(16, 16), block_size=(8, 8), periodic=(False, False))
    with pytest.raises(ValueError, match="periodic"):
        tpamr.NavierStokesSystem(walls, dynamic_viscosity=0.01)
    mixed = tpamr.BlockTreeMesh(base_mesh=(16, 16), block_size=(8, 8), max_level=1,
                                refine=lambda k, o, e: k.i == 0 and k.j == 0)
    with pytest.raises(ValueError, match="uniform"):
        tpamr.NavierStokesSystem(mixed, dynamic_viscosity=0.01)


def test_staggered_multigrid_levels():
    """'staggered base-mesh multigrid' (test/operators/staggered_diffusion.jl:
    1752-1766), periodic: base meshes (16,16), (8,8), (4,4) with coarse_size
    16; a single root block keeps one level (:1694-1736); leaf levels are
    pruned before the base mesh is halved; the Taylor-Green mesh."""
    cases = [
        (dict(base_mesh=(16, 16), block_size=(4, 4)), 16, [256, 64, 16]),
        (dict(base_mesh=(16, 16), block_size=(16, 16)), 64, [256]),
        (MESHES["refined_4"], 64, [1024, 256, 64]),
        (MESHES["taylor_green_128"], 64, [16384, 4096, 1024, 256]),
    ]
    for kw, coarse, want in cases:
        pm = tpamr.BlockTreeMesh(**dict(kw, periodic=(True, True)))
        sysg = tpamr.NavierStokesSystem(pm, dynamic_viscosity=0.2, coarse_size=coarse)
        assert sysg.viscous_levels == want, (kw, sysg.viscous_levels)
        sysg.close()


@pytest.mark.parametrize("name", ["refined_4", "roots_8", "taylor_green_128"])
def test_multigrid_viscous_solve_matches_oracle(name):
    """The viscous solve preconditioned by the staggered multigrid: the same
    iteration counts per component as the oracle's restatement of the
    reference's solver, and the same solution (1e-12)."""
    om, pm = pair(name)
    nu, dt = 0.5, 0.02
    Lx, Ly = MESHES[name]["extent"]
    old = sample(om, lambda x, y: np.sin(2 * np.pi * x / Lx) * np.cos(4 * np.pi * y / Ly),
                 lambda x, y: np.cos(2 * np.pi * x / Lx))
    conv = old + 0.05 * np.sin(np.arange(om.nfaces) * 0.37)
    sysg = tpamr.NavierStokesSystem(pm, dynamic_viscosity=nu)
    dst = old.copy()
    sx, sy = sysg.implicit_midpoint_diffusion(dst, conv, old, dt, tolerance=1e-11)
    want, st = implicit_midpoint_mg(StaggeredDiffusion(om, nu), mg_params(name), conv, old, old, dt, tol=1e-11)
    assert (sx.cycles, sy.cycles) == (st[0]["cycles"], st[1]["cycles"])
    assert sx.cycles >= 2
    assert rel(dst, want) < 1e-12
    sysg.close()


@pytest.mark.parametrize("base, want_levels", [((96, 96), [9216, 2304]), ((48, 80), [3840])])
def test_viscous_solve_with_large_coarsest_level_matches_oracle(base, want_levels):
    """Root layouts whose staggered multigrid stops above 1024 unknowns per
    component: 6x6 roots of 16^2 cells halve once to 3x3 (coarsest 2304),
    3x5 roots do not halve (one level of 3840). The reference solves the
    coarsest level exactly at any size (dense or sparse LU,
    implicit_diffusion.jl:343-347); the library by the periodic convolution
    with the first column of the circulant inverse. The oracle's exact coarse
    LU gives the same iteration counts per component and the same solution
    (1e-12)."""
    kw = dict(base_mesh=base, block_size=(16, 16), extent=(1.0, 1.0), periodic=(True, True))
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    nu, dt = 0.5, 0.02
    old = sample(om, lambda x, y: np.sin(2 * np.pi * x) * np.cos(4 * np.pi * y), lambda x, y: np.cos(2 * np.pi * x))
    conv = old + 0.05 * np.sin(np.arange(om.nfaces) * 0.37)
    sysg = tpamr.NavierStokesSystem(pm, dynamic_viscosity=nu)
    assert sysg.viscous_levels == want_levels
    dst = old.copy()
    sx, sy = sysg.implicit_midpoint_diffusion(dst, conv, old, dt, tolerance=1e-11)
    want, st = implicit_midpoint_mg(StaggeredDiffusion(om, nu), (base, (16, 16), 0), conv, old, old, dt, tol=1e-11)
    assert (sx.cycles, sy.cycles) == (st[0]["cycles"], st[1]["cycles"])
    assert sx.converged and sy.converged
    assert rel(dst, want) < 1e-12
    sysg.close()


def test_implicit_midpoint_amplification_on_one_root():
    """'implicit midpoint staggered elliptic solve'
    (test/operators/staggered_diffusion.jl:1694-1728): a 16^2 root block
    (one level: the dense coarse solve), viscosity 0.1, dt 0.1, a
    sin(2 pi x) / cos(2 pi y) field, default tolerance: the discrete
    amplification factor to 2e-13."""
    kw = dict(block_size=(16, 16), periodic=(True, True))
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    nu, dt = 0.1, 0.1
    sysg = tpamr.NavierStokesSystem(pm, dynamic_viscosity=nu)
    old = sample(om, lambda x, y: np.sin(2 * np.pi * x), lambda x, y: np.cos(2 * np.pi * y))
    out = np.zeros(om.nfaces)
    sx, sy = sysg.implicit_midpoint_diffusion(out, old, old, dt)
    dx = 1.0 / 16
    lam = -4 * np.sin(np.pi / 16) ** 2 / dx ** 2
    g = (1 + dt * nu * lam / 2) / (1 - dt * nu * lam / 2)
    assert np.max(np.abs(out - g * old)) <= 2e-13 + 2e-13 * np.max(np.abs(g * old))
    assert sx.converged and sy.converged and sysg.viscous_levels == [256]
    sysg.close()


def test_taylor_green_at_1024_properties():
    """At 1024^2 cells (beyond the oracle's reach per step) two steps keep the
    size-independent properties: the projected velocity's divergence within
    the solve's tolerance, the fields within the second-order truncation of
    the exact decaying Taylor-Green solution, and the kinetic energy decaying
    at the viscous rate exp(-4 nu t) (its transport error is of higher order)."""
    kw = dict(base_mesh=(1024, 1024), block_size=(16, 16), extent=(2 * np.pi, 2 * np.pi), periodic=(True, True))
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    nu, dt = 0.02, 1e-3
    sysg = tpamr.NavierStokesSystem(pm, dynamic_viscosity=nu)
    u, p = tg(om), np.zeros(pm.ncells)
    e0 = np.sum(u ** 2)
    for _ in range(2):
        st = tpamr.step(u, p, sysg, dt, tolerance=1e-12)
        assert st.converged
    vx, vy = sysg.viscous_stats()
    assert vx.converged and vy.converged
    sysg.close()
    h = 2 * np.pi / 1024
    assert np.max(np.abs(om.divergence(u))) <= 1e-8
    assert rel(u, tg(om, 2 * dt, nu)) <= 4 * h * h
    assert abs(np.sum(u ** 2) / e0 - np.exp(-4 * nu * 2 * dt)) <= 1e-5
