"""GPU parity: the HIP path (through the C ABI) against the CPU FP64 oracle.

Tolerances: operator applications and coefficient kernels agree to round-off
(1e-13 relative); an FP64 V-cycle to 1e-11 (its component-mean removal is
applied once at the end of the cycle instead of after every sweep, which only
moves round-off); FP64 solves to 1e-10 relative with identical iteration
counts; FP32-V-cycle solves and FP32 refinement solves (FP64 iterative
refinement around an inner CG with FP32 vectors) reach the same FP64 residual
target and their solution is compared with the FP64 oracle solution.
"""
import math
import os
import sys

import numpy as np
import pytest

from cases import MESHES, DROP_CENTRE, DROP_RADIUS, disk_fraction, techref_noise

pytestmark = pytest.mark.gpu

import tpamr  # noqa: E402
from oracle import OMesh, OSolver  # noqa: E402


def _print(*a):
    print(*a, flush=True)


_CACHE = {}


def meshes(name):
    if name not in _CACHE:
        kw = MESHES[name]
        _CACHE[name] = (OMesh(**kw), tpamr.BlockTreeMesh(**kw))
    return _CACHE[name]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def kappa_for(pm, kind):
    if kind == "one":
        return np.ones(pm.nfaces)
    if kind == "smooth":
        return np.exp(0.4 * np.sin(np.arange(1, pm.nfaces + 1)))
    if kind == "drop":
        x, y = pm.cell_centers()
        dx, dy = pm.cell_spacing()
        c = disk_fraction(x, y, dx, dy, DROP_CENTRE, DROP_RADIUS)
        om, _ = None, None
        return c
    raise ValueError(kind)


CASES = [("uniform64", "one"), ("amr32", "one"), ("amr_mixed", "smooth"), ("pruned_tree", "smooth"),
         ("periodic_tg", "one"), ("circle16", "smooth"), ("amr_block4", "smooth"), ("amr_block32", "smooth")]


@pytest.fixture(scope="module", params=["fp64", "fp32_vcycle", "fp32_refine"])
def precision(request):
    return request.param


@pytest.mark.parametrize("name,kind", CASES)
def test_apply_matches_oracle(name, kind):
    _print("case apply", name, kind)
    om, pm = meshes(name)
    k = kappa_for(pm, kind)
    os_ = OSolver(om, k)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k))
    rng = np.random.default_rng(7)
    x = rng.standard_normal(pm.ncells)
    assert rel(gs.apply(x), os_.apply(x)) < 1e-13
    gs.close()


@pytest.mark.parametrize("name,kind", CASES)
def test_vcycle_matches_oracle(name, kind, precision):
    _print("case vcycle", name, kind, precision)
    om, pm = meshes(name)
    k = kappa_for(pm, kind)
    os_ = OSolver(om, k)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), precision=precision)
    rng = np.random.default_rng(11)
    r = rng.standard_normal(pm.ncells)
    r -= pm.volumes * r.sum() / pm.volumes.sum()
    zo = os_.vcycle(r)
    zg = gs.ldiv(r)
    tol = 1e-11 if precision == "fp64" else 2e-5
    assert rel(zg, zo) < tol
    # the V-cycle output has zero volume-weighted mean
    assert abs(np.dot(pm.volumes, zg)) <= 1e-12 * np.sum(pm.volumes * np.abs(zg))
    gs.close()


@pytest.mark.parametrize("name,kind", CASES)
def test_solve_matches_oracle(name, kind, precision):
    _print("case solve", name, kind, precision)
    om, pm = meshes(name)
    k = kappa_for(pm, kind)
    os_ = OSolver(om, k)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), precision=precision)
    b = techref_noise(pm.ncells)
    xo, so = os_.solve(b, rtol=1e-10, max_cycles=200)
    xg = np.zeros(pm.ncells)
    sg = gs.solve(xg, b, rtol=1e-10, max_cycles=200)
    assert sg.converged and so["converged"]
    if precision == "fp64":
        assert sg.cycles == so["cycles"]
        assert rel(xg, xo) < 1e-10
        assert abs(sg.initial_linf - so["initial_linf"]) <= 1e-12 * so["initial_linf"]
    elif precision == "fp32_vcycle":
        assert abs(sg.cycles - so["cycles"]) <= 2
        # both stop at the same strong-residual target; the solutions then
        # differ by at most that residual through the operator's inverse
        assert rel(xg, xo) < 1e-7
    else:
        # each refinement step restarts the inner Krylov recurrence (1e-5
        # reduction per step): a few more iterations than one FP64 run
        assert sg.cycles <= so["cycles"] + 2 + so["cycles"] // 3
        assert rel(xg, xo) < 1e-7
    assert sg.final_linf <= sg.target_linf
    gs.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_vcycle"])
@pytest.mark.parametrize("name,kind", [("amr_mixed", "smooth"), ("circle16", "smooth")])
def test_krylov_seam_matches_solve(name, kind, precision):
    """tpamr_krylov (the _run_krylov! seam) from x0 = 0 on the integrated,
    compatible rhs reproduces the first Krylov run of the full solve."""
    _print("case krylov seam", name, kind, precision)
    om, pm = meshes(name)
    k = kappa_for(pm, kind)
    os_ = OSolver(om, k)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), precision=precision)
    b = techref_noise(pm.ncells)
    xs = np.zeros(pm.ncells)
    st = gs.solve(xs, b, rtol=1e-10, max_cycles=200)
    V = pm.volumes
    bi = V * b
    bi -= V * (bi.sum() / V.sum())
    x, it, reason = gs.run_krylov(bi, np.zeros(pm.ncells), st.target_linf, 200)
    assert reason == "strong residual target"
    x -= np.dot(V, x) / V.sum()
    assert it == st.cycles
    assert rel(x, xs) < 1e-12
    # the oracle's operator confirms the strong residual of the iterate
    r = (bi - os_.apply(x)) / V
    assert np.max(np.abs(r)) <= 1.01 * st.target_linf
    with pytest.raises(ValueError):
        gs.run_krylov(bi, np.zeros(pm.ncells), -1.0)
    gs.close()


TECHREF = [("uniform64", 4096, 4, 10), ("uniform128", 16384, 5, 10), ("amr32", 7552, 5, 10), ("amr64", 19648, 6, 11)]


@pytest.mark.parametrize("name,unknowns,levels,cycles", TECHREF)
def test_published_iteration_counts_on_gpu(name, unknowns, levels, cycles):
    """docs/techref/figures/generated/checks.tex:200-229 of the reference."""
    _print("case techref", name)
    _, pm = meshes(name)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, 1.0))
    assert pm.ncells == unknowns and len(gs.levels) == levels
    x = np.zeros(pm.ncells)
    st = gs.solve(x, techref_noise(pm.ncells), rtol=1e-10, max_cycles=20000)
    assert st.converged and st.cycles == cycles
    factor = (st.final_l2 / st.initial_l2) ** (1 / st.cycles)
    assert round(factor, 2) in (0.08, 0.09)
    gs.close()


@pytest.mark.parametrize("name", ["uniform128", "amr64"])
def test_drop_density_ratio_counts_on_gpu(name):
    """drop 1000:1 rows of checks.tex:230-241: 15 and 16 PCG iterations."""
    _print("case drop", name)
    om, pm = meshes(name)
    x, y = pm.cell_centers()
    dx, dy = pm.cell_spacing()
    c = disk_fraction(x, y, dx, dy, DROP_CENTRE, DROP_RADIUS)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, 1.0))
    k = tpamr.mixture_face_coefficients(gs, c, (1000.0, 1.0))
    # every product and sum rounded on its own, as the reference's: bit-exact
    assert np.array_equal(k, om.mixture_coefficients(c, 1000.0, 1.0))
    cn = c.copy()
    cn[::97] = np.nan
    kn = tpamr.mixture_face_coefficients(gs, cn, (1000.0, 1.0))
    ko = om.mixture_coefficients(cn, 1000.0, 1.0)
    assert np.isnan(kn).any() and np.array_equal(np.isnan(kn), np.isnan(ko))
    assert np.array_equal(kn[~np.isnan(kn)], ko[~np.isnan(ko)])
    tpamr.refresh_two_phase(gs, c, (1000.0, 1.0))
    xs = np.zeros(pm.ncells)
    st = gs.solve(xs, techref_noise(pm.ncells), rtol=1e-10, max_cycles=20000)
    assert st.converged and st.cycles == {"uniform128": 15, "amr64": 16}[name]
    gs.close()


@pytest.mark.parametrize("name", ["uniform128", "amr64"])
def test_drop_density_ratio_refinement(name):
    """FP32 refinement on the 1000:1 drop: the FP64 strong-residual target is
    met and the pressure matches the FP64 GPU solve of the same system."""
    _print("case drop refinement", name)
    _, pm = meshes(name)
    x, y = pm.cell_centers()
    dx, dy = pm.cell_spacing()
    c = disk_fraction(x, y, dx, dy, DROP_CENTRE, DROP_RADIUS)
    sols = {}
    for prec in ("fp64", "fp32_refine"):
        gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, 1.0), precision=prec)
        tpamr.refresh_two_phase(gs, c, (1000.0, 1.0))
        xs = np.zeros(pm.ncells)
        st = gs.solve(xs, techref_noise(pm.ncells), rtol=1e-10, max_cycles=20000)
        assert st.converged and st.final_linf <= st.target_linf
        sols[prec] = (xs, st)
        gs.close()
    (x64, s64), (x32, s32) = sols["fp64"], sols["fp32_refine"]
    _print(f"fp64 {s64.cycles} iterations, fp32_refine {s32.cycles} iterations")
    assert s32.cycles <= s64.cycles + 2 + s64.cycles // 3
    assert rel(x32, x64) < 1e-7


def test_two_phase_face_kernels_match_oracle():
    _print("case two-phase kernels")
    om, pm = meshes("amr64")
    x, y = pm.cell_centers()
    dx, dy = pm.cell_spacing()
    c = disk_fraction(x, y, dx, dy, DROP_CENTRE, DROP_RADIUS)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, 1.0))
    k = tpamr.mixture_face_coefficients(gs, c, (1000.0, 1.0))
    tpamr.refresh_two_phase(gs, c, (1000.0, 1.0))
    os_ = OSolver(om, k)
    curv = np.full(pm.nfaces, 5.0)
    curv[::97] = np.nan
    jg = tpamr.surface_tension_jumps(gs, c, curv, sigma=0.072)
    jo = om.cgfm_jumps(c, curv, 0.072)
    assert np.array_equal(jg, jo)
    jc = tpamr.surface_tension_jumps(gs, c, curv, sigma=0.072, formulation=tpamr.CSF)
    assert np.array_equal(jc, om.cgfm_jumps(c, curv, 0.072, csf=True))
    assert rel(tpamr.weighted_jump_force(gs, jg), os_.weighted_jump_force(jo)) < 1e-13
    p = np.sin(3 * x) * np.cos(2 * y)
    assert rel(tpamr.weighted_gradient(gs, p), os_.weighted_gradient(p)) < 1e-13
    u = np.cos(np.arange(pm.nfaces) * 0.37)
    assert rel(tpamr.divergence(gs, u), om.divergence(u)) < 1e-13
    gs.close()


def test_projection_matches_oracle():
    _print("case projection")
    om, pm = meshes("amr_mixed")
    k = kappa_for(pm, "smooth")
    os_ = OSolver(om, k)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k))
    u0 = np.cos(np.arange(pm.nfaces) * 0.61)
    p0 = np.zeros(pm.ncells)
    dt = 1e-3
    uo, po, phio, so = os_.project(u0, p0, dt, tolerance=1e-12, max_cycles=200)
    ug, pg = u0.copy(), p0.copy()
    sg, phig = tpamr.project(ug, pg, gs, dt, tolerance=1e-12, max_cycles=200)
    assert sg.converged and sg.cycles == so["cycles"]
    assert rel(phig, phio) < 1e-10
    assert rel(ug, uo) < 1e-10
    assert rel(pg, po) < 1e-10
    gs.close()


def test_adjoint_matches_oracle_and_finite_differences():
    _print("case adjoint")
    om, pm = meshes("circle16")
    k = kappa_for(pm, "smooth")
    os_ = OSolver(om, k)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k))
    rng = np.random.default_rng(3)
    b = techref_noise(pm.ncells)
    w = rng.standard_normal(pm.ncells)
    p, pull = tpamr.pressure_solve_rrule(gs, b, rtol=1e-13)
    db, dk = pull(w)
    po, _ = os_.solve(b, rtol=1e-13, max_cycles=200)
    dbo, dko, _, _ = os_.adjoint(w, po, rtol=1e-13)
    assert rel(db, dbo) < 1e-9 and rel(dk, dko) < 1e-9
    # central differences along random directions (J = w . p)
    eps = 1e-6
    for q in range(3):
        d = rng.standard_normal(pm.ncells)
        jp = w @ tpamr.pressure_solve(gs, b + eps * d, rtol=1e-13)[0]
        jm = w @ tpamr.pressure_solve(gs, b - eps * d, rtol=1e-13)[0]
        fd = (jp - jm) / (2 * eps)
        assert abs(fd - db @ d) <= 1e-6 * max(abs(fd), 1e-30)
    for q in range(2):
        dkap = rng.standard_normal(pm.nfaces) * k
        vals = []
        for sgn in (1, -1):
            gs.operator.replace_face_coefficient(k + sgn * eps * dkap)
            vals.append(w @ tpamr.pressure_solve(gs, b, rtol=1e-13)[0])
        gs.operator.replace_face_coefficient(k)
        fd = (vals[0] - vals[1]) / (2 * eps)
        assert abs(fd - dk @ dkap) <= 1e-6 * max(abs(fd), 1e-30)
    gs.close()


def test_adjoint_two_phase_circle_twenty_directions():
    """Differentiation case on the two-phase circle mesh (one 16x16 root, levels
    3-7 within 4 finest cells of r = 0.15 at (0.5, 0.6), density 1000:1): J = w.p
    with seeded w ~ N(0,1); the rrule gradient w.r.t. the rhs and the face
    coefficients beta_f = 1/rho_f against the FP64 CPU (oracle) gradient and
    against central differences (eps = 1e-6) along 20 seeded directions.

    At 1000:1 a cold solve stops at the round-off floor, whose error divided
    by 2 eps would swamp the difference quotient. Each perturbed solution is
    therefore p0 + dp, one step of iterative refinement from the unperturbed
    solution p0 (A dp = b - A p0 for the perturbed A or b), and the common
    w.p0 is cancelled before rounding: J+ - J- = w.(dp+ - dp-)."""
    from tpamr import workloads as W
    _print("case adjoint two-phase circle")
    mesh, circ = W.single_circle_mesh()
    om = OMesh(base_mesh=mesh.base_mesh, block_size=mesh.block_size, periodic=mesh.periodic,
               max_level=mesh.max_level, refined_keys=mesh.refined_keys)
    c = W.liquid_fraction(mesh, circ)
    k = om.mixture_coefficients(c, 1000.0, 1.0)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(mesh, k))
    V = mesh.volumes
    b = -om.divergence(W.smooth_velocity(mesh))
    b -= np.dot(V, b) / V.sum()
    rng = W.SplitMix64(W.SEED + 4)
    w = np.array([rng.normal() for _ in range(mesh.ncells)])
    p0, pull = tpamr.pressure_solve_rrule(gs, b, rtol=1e-13)
    db, dk = pull(w)
    os_ = OSolver(om, k)
    po, _ = os_.solve(b, rtol=1e-13, max_cycles=200)
    dbo, dko, _, _ = os_.adjoint(w, po, rtol=1e-13)
    assert rel(p0, po) < 1e-10
    assert rel(db, dbo) < 1e-9 and rel(dk, dko) < 1e-9

    def correction(rhs):
        r = rhs - gs.apply(p0) / V
        r -= np.dot(V, r) / V.sum()
        dp = np.zeros(mesh.ncells)
        gs.solve(dp, r, rtol=1e-12, max_cycles=200)
        return dp

    eps = 1e-6
    worst = 0.0
    for q in range(20):
        if q % 2 == 0:
            d = np.array([rng.normal() for _ in range(mesh.ncells)])
            d -= np.dot(V, d) / V.sum()
            dpp, dpm = correction(b + eps * d), correction(b - eps * d)
            ad = db @ d
        else:
            d = np.array([rng.normal() for _ in range(mesh.nfaces)]) * k
            vals = []
            for sgn in (1, -1):
                gs.operator.replace_face_coefficient(k + sgn * eps * d)
                vals.append(correction(b))
            gs.operator.replace_face_coefficient(k)
            dpp, dpm = vals
            ad = dk @ d
        fd = w @ (dpp - dpm) / (2 * eps)
        err = abs(fd - ad) / max(abs(fd), 1e-30)
        worst = max(worst, err)
        _print(f"direction {q} ({'rhs' if q % 2 == 0 else 'beta'}): fd {fd:.12e} adjoint {ad:.12e} rel {err:.2e}")
        assert err <= 1e-6, (q, fd, ad)
    _print(f"adjoint vs central differences: worst relative difference {worst:.2e} over 20 directions")
    gs.close()


def test_errors_are_reported():
    _print("case errors")
    _, pm = meshes("uniform64")
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, 1.0))
    x = np.zeros(pm.ncells)
    with pytest.raises(ValueError):
        gs.solve(x, np.full(pm.ncells, np.nan))
    with pytest.raises(ValueError):
        gs.solve(x, np.ones(pm.ncells), compatibility="error")
    with pytest.raises(tpamr.DimensionMismatch):
        gs.solve(x, np.ones(pm.ncells - 1))
    st = gs.solve(x, techref_noise(pm.ncells), max_cycles=0)
    assert not st.converged and st.cycles == 0
    gs.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_vcycle", "fp32_refine"])
@pytest.mark.parametrize("name", ["amr_mixed", "circle16", "pruned_tree"])
def test_device_refresh_matches_host_refresh(name, precision):
    """A coefficient refresh after the first runs on the device (restriction,
    conductances, irregular rows, Jacobi diagonal and weights, FP32 copies); it
    performs the host pass's operations in the same order, so a solve after it
    is bit for bit the solve after a host refresh."""
    _, pm = meshes(name)
    k1 = kappa_for(pm, "smooth")
    k2 = np.exp(0.7 * np.cos(0.37 * np.arange(1, pm.nfaces + 1)))
    x, y = pm.cell_centers()
    rhs = np.sin(3 * x) * np.cos(5 * y) + 0.1 * np.cos(7 * x * y)
    out = {}
    for host in (True, False):
        if host:
            os.environ["TPAMR_HOST_REFRESH"] = "1"
        try:
            s = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k1), precision=precision, device=0)
        finally:
            os.environ.pop("TPAMR_HOST_REFRESH", None)
        s.operator.replace_face_coefficient(k2)
        s.refresh()
        p = np.zeros(pm.ncells)
        st = s.solve(p, rhs, rtol=1e-10)
        nf1 = s.level_info(1)["nfaces"]
        k_l1 = np.zeros(nf1)
        tpamr._lib.check(tpamr.lib().tpamr_level_kappa(s.ctx.h, 1, tpamr._lib.ptr(k_l1)), s.ctx.h)
        out[host] = (p, st.cycles, [lv["omega"] for lv in s.levels], k_l1)
        if not host:
            # host copies of the level values follow on demand
            assert np.array_equal(k_l1, out[True][3])
            # a negative coefficient is refused by the device pass before it
            # changes any level
            kneg = np.where(np.arange(pm.nfaces) == 3, -1.0, k2)
            with pytest.raises(ValueError, match="nonnegative"):
                tpamr._lib.check(tpamr.lib().tpamr_set_face_coefficients(s.ctx.h, tpamr._lib.ptr(kneg)), s.ctx.h)
            # a zero one takes the host refresh (zeros may split the domain
            # into components): the solve agrees with a fresh solver's (the
            # refreshed coarse inverse may come from a Newton-Schulz update)
            kzero = np.where(np.arange(pm.nfaces) == 3, 0.0, k2)
            s.operator.replace_face_coefficient(kzero)
            pz = np.zeros(pm.ncells)
            sz = s.solve(pz, rhs, rtol=1e-10)
            fresh = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, kzero), precision=precision, device=0)
            pf = np.zeros(pm.ncells)
            fresh.solve(pf, rhs, rtol=1e-10)
            fresh.close()
            assert sz.converged and sz.removed_rhs_component_anchors == [1]
            assert np.max(np.abs(pz - pf)) <= 1e-8 * np.max(np.abs(pf))
            s.operator.replace_face_coefficient(k2)
            s.refresh()
            q = np.zeros(pm.ncells)
            s.solve(q, rhs, rtol=1e-10)
            assert np.max(np.abs(q - p)) <= 1e-8 * np.max(np.abs(p))
        s.close()
    _print(name, precision, out[True][1], out[False][1])
    assert out[True][1] == out[False][1]
    assert out[True][2] == out[False][2]
    assert np.array_equal(out[True][0], out[False][0])


def test_projection_with_inherited_divergence_matches_oracle():
    """project! on a shortened step (dt < reference dt) keeps part of the
    inherited divergence (projection.jl:130-156); stats.compatible follows
    _projection_compatibility! (:100-127)."""
    om, pm = meshes("amr_mixed")
    k = kappa_for(pm, "smooth")
    os_ = OSolver(om, k)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k))
    u0 = np.cos(np.arange(pm.nfaces) * 0.61)
    inh = om.divergence(np.sin(np.arange(pm.nfaces) * 0.37))
    dt, tau = 1e-3, 2.5e-3
    uo, po, phio, so = os_.project(u0, np.zeros(pm.ncells), dt, tolerance=1e-12, max_cycles=200, inherited=inh,
                                   reference_dt=tau)
    ug, pg = u0.copy(), np.zeros(pm.ncells)
    sg, phig = tpamr.project(ug, pg, gs, dt, tolerance=1e-12, max_cycles=200, inherited_divergence=inh,
                             reference_dt=tau)
    assert sg.converged and sg.cycles == so["cycles"] and sg.compatible == bool(so["compatible"])
    assert rel(phig, phio) < 1e-10 and rel(ug, uo) < 1e-10
    # a constant inherited divergence cannot be projected out
    _, _, _, sc = os_.project(u0, np.zeros(pm.ncells), dt, tolerance=1e-12, inherited=np.ones(pm.ncells),
                              reference_dt=tau)
    sgc, _ = tpamr.project(u0.copy(), np.zeros(pm.ncells), gs, dt, tolerance=1e-12,
                           inherited_divergence=np.ones(pm.ncells), reference_dt=tau)
    assert not sc["compatible"] and not sgc.compatible
    with pytest.raises(ValueError):
        tpamr.project(u0.copy(), np.zeros(pm.ncells), gs, dt, reference_dt=0.5 * dt)
    gs.close()


def test_pressure_predictor_matches_oracle():
    """_pressure_predictor! (single_phase/stepping.jl:144-168): u -= dt G p."""
    om, pm = meshes("amr_mixed")
    k = kappa_for(pm, "smooth")
    os_ = OSolver(om, k)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k))
    x, y = pm.cell_centers()
    p = np.sin(3 * x) * np.cos(2 * y)
    u0 = np.cos(np.arange(pm.nfaces) * 0.61)
    ug = tpamr.pressure_predictor(u0.copy(), p, gs, 1e-3)
    uo = u0 - 1e-3 * os_.weighted_gradient(p)
    assert rel(ug, uo) < 1e-13
    gs.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_coarse_newton_schulz_update_matches_gauss_jordan(precision):
    """After the first refresh the dense coarse inverse is updated from the
    previous one by Newton-Schulz steps (GEMMs); a Gauss-Jordan inverse at
    every refresh (TPAMR_COARSE_GJ=1) gives the same solves."""
    _, pm = meshes("pruned_tree")
    k1 = kappa_for(pm, "smooth")
    x, y = pm.cell_centers()
    rhs = np.sin(3 * x) * np.cos(5 * y)
    # small changes (Newton-Schulz steps), then a 1000:1 jump (the update
    # diverges: Gauss-Jordan), then a small change from there
    jump = np.where(x[pm.face_minus] < 0.4, 1.0, 1e-3) * k1
    seq = [k1 * (1 + 0.01 * (r + 1) * np.cos(np.arange(pm.nfaces))) for r in range(3)] + [jump, jump * 1.01]
    out = []
    for exact in (True, False):
        if exact:
            os.environ["TPAMR_COARSE_GJ"] = "1"
        try:
            s = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k1), precision=precision, device=0)
        finally:
            os.environ.pop("TPAMR_COARSE_GJ", None)
        sols = []
        for kk in seq:
            s.operator.replace_face_coefficient(kk)
            s.refresh()
            p = np.zeros(pm.ncells)
            st = s.solve(p, rhs, rtol=1e-10)
            sols.append((p, st.cycles))
        out.append(sols)
        s.close()
    for (pg, cg), (pn, cn) in zip(*out):
        assert cg == cn
        assert rel(pn, pg) < 1e-12


def test_solvers_built_and_closed_repeatedly_release_device_memory():
    """A regridding run builds a new pressure solver at every regrid
    (`_rebuild_pressure_solver!`, src/navier_stokes/workspace.jl:82-112):
    building, refreshing (device refresh, Newton-Schulz coarse update),
    solving and closing solvers again and again leaves the device's free
    memory where it was."""
    import torch
    from tpamr import workloads as W
    pm = W.circles_mesh(levels=(3, 7), n=16)[0]
    x, y = pm.cell_centers()
    rhs = np.sin(3 * x) * np.cos(5 * y)
    k = np.exp(0.3 * np.sin(np.arange(pm.nfaces)))

    def cycle():
        for precision in ("fp64", "fp32_refine"):
            s = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), precision=precision, device=0)
            for r in range(3):
                s.operator.replace_face_coefficient(k * (1 + 0.01 * r))
                s.refresh()
                p = np.zeros(pm.ncells)
                assert s.solve(p, rhs, rtol=1e-10).converged
            s.close()
        torch.cuda.synchronize()
        return torch.cuda.mem_get_info()[0]

    cycle()  # one-time runtime and library state
    free0 = cycle()
    for _ in range(3):
        free = cycle()
    print("free device memory after the second and the fifth cycle:", free0, free, flush=True)
    assert abs(free - free0) <= 16 << 20
