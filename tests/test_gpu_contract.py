"""GPU: the solve contract around the kernels, through the C ABI, against the
CPU FP64 oracle.

- One context across meshes of different sizes: the device coefficient
  refresh follows the current mesh (the refresh plan is rebuilt with it).
- tpamr_solve_device equals tpamr_solve bit for bit, in every precision mode,
  and rejects non-finite device inputs.
- The stopping statistics (initial and final strong residuals, target with
  the guarded round-off floor, removed rhs mean) match the oracle's
  `_solve_krylov_with_stats!` restatement, which is pinned to the
  reference's own known answers in tests/test_oracle_stopping.py.
- The periodic nullspace convention (test/operators/poisson.jl:921-933).
- The Julia shim's `solve!` override (Seam 2), mirrored call for call.
"""
import ctypes as C
import math

import numpy as np
import pytest

from cases import MESHES, DROP_CENTRE, DROP_RADIUS, disk_fraction, techref_noise

pytestmark = pytest.mark.gpu

import tpamr  # noqa: E402
from tpamr import _lib as L  # noqa: E402
from oracle import OMesh, OSolver  # noqa: E402

MODES = {"fp64": L.TPAMR_FP64, "fp32_vcycle": L.TPAMR_FP32_VCYCLE, "fp32_refine": L.TPAMR_FP32_REFINE}


def _print(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def drop_kappa(om, pm, centre=DROP_CENTRE, radius=DROP_RADIUS):
    x, y = pm.cell_centers()
    dx, dy = pm.cell_spacing()
    c = disk_fraction(x, y, dx, dy, centre, radius)
    return om.mixture_coefficients(c, 1000.0, 1.0)


def raw_solve(ctx, b, x, rtol, max_cycles=200, compat=0, atol=0.0):
    st = L.Stats()
    L.check(L.lib().tpamr_solve(ctx.h, L.ptr(b), L.ptr(x), float(rtol), float(atol), int(max_cycles), int(compat),
                                C.byref(st)), ctx.h)
    return st


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_one_context_across_meshes(precision):
    """One context, meshes attached in turn (larger, smaller, larger): each
    gets a host refresh (hierarchy build) and then a device refresh (the plan
    of the current mesh), and its solve matches the oracle."""
    ctx = L.Context(0, MODES[precision])
    try:
        for name in ("amr64", "amr32", "circle16", "amr64"):
            _print("case remesh", precision, name)
            kw = MESHES[name]
            om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
            pm.attach(ctx)
            k = drop_kappa(om, pm)
            k_first = k * np.exp(0.3 * np.sin(np.arange(1, pm.nfaces + 1)))
            for kk in (k_first, k, k):  # build + host refresh, then two device refreshes
                L.check(L.lib().tpamr_set_face_coefficients(ctx.h, L.ptr(np.ascontiguousarray(kk))), ctx.h)
            ns, os_start = C.c_int(), C.c_int()
            L.check(L.lib().tpamr_hierarchy_reuse(ctx.h, C.byref(ns), C.byref(os_start)), ctx.h)
            # amr32 after amr64 takes over the previous agglomerated levels
            # 1024, 256 and 64 (and the coarse inverse as a Newton-Schulz start)
            assert (ns.value, os_start.value) == ((2, 3) if name == "amr32" else (-1, -1))
            b = techref_noise(pm.ncells)
            x = np.zeros(pm.ncells)
            st = raw_solve(ctx, b, x, 1e-10)
            xo, so = OSolver(om, k).solve(b, rtol=1e-10, max_cycles=200)
            assert st.converged and so["converged"]
            if precision == "fp64":
                assert st.cycles == so["cycles"]
                assert rel(x, xo) < 1e-10
            else:
                assert rel(x, xo) < 1e-7
    finally:
        ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_vcycle", "fp32_refine"])
def test_solve_device_equals_solve(precision):
    """tpamr_solve_device (the entry point the benchmark times) gives the
    host entry point's result bit for bit, and both reject non-finite input
    without disturbing the next solve."""
    import torch
    _print("case solve_device", precision)
    om, pm = OMesh(**MESHES["amr64"]), tpamr.BlockTreeMesh(**MESHES["amr64"])
    k = drop_kappa(om, pm)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), precision=precision)
    b = techref_noise(pm.ncells)
    x0 = 0.01 * np.cos(np.arange(pm.ncells))
    xh = x0.copy()
    sh = gs.solve(xh, b, rtol=1e-10, max_cycles=200)
    db = torch.tensor(b, dtype=torch.float64, device="cuda")
    dx = torch.tensor(x0, dtype=torch.float64, device="cuda")
    sd = gs.solve_device(db.data_ptr(), dx.data_ptr(), rtol=1e-10, max_cycles=200)
    torch.cuda.synchronize()
    assert sh.converged and sd.converged
    assert sd.cycles == sh.cycles
    assert np.array_equal(dx.cpu().numpy(), xh)
    for f in ("initial_linf", "final_linf", "target_linf", "removed_rhs_mean", "initial_l2", "final_l2"):
        assert getattr(sd, f) == getattr(sh, f), f
    # the device rhs is not modified
    assert np.array_equal(db.cpu().numpy(), b)
    bad = db.clone()
    bad[123] = float("nan")
    with pytest.raises(ValueError, match="finite"):
        gs.solve_device(bad.data_ptr(), dx.data_ptr(), rtol=1e-10)
    badx = dx.clone()
    badx[7] = float("inf")
    with pytest.raises(ValueError, match="finite"):
        gs.solve_device(db.data_ptr(), badx.data_ptr(), rtol=1e-10)
    dx2 = torch.tensor(x0, dtype=torch.float64, device="cuda")
    sd2 = gs.solve_device(db.data_ptr(), dx2.data_ptr(), rtol=1e-10, max_cycles=200)
    torch.cuda.synchronize()
    assert sd2.cycles == sh.cycles and np.array_equal(dx2.cpu().numpy(), xh)
    gs.close()


STATS_CASES = [("amr64", DROP_CENTRE, DROP_RADIUS), ("circle16", (0.5, 0.6), 0.15)]


@pytest.mark.parametrize("precision", ["fp64", "fp32_vcycle", "fp32_refine"])
@pytest.mark.parametrize("name,centre,radius", STATS_CASES)
def test_stopping_statistics_match_oracle(name, centre, radius, precision):
    """1000:1 drop and circle meshes, rtol 1e-10, from a nonzero initial
    pressure: initial norms, removed rhs mean and target equal the oracle's
    (the target is rtol times the initial strong residual); the final strong
    residual meets the target, and in FP64 equals the oracle's to the
    accuracy of the iterates' last digits."""
    _print("case stats", name, precision)
    om, pm = OMesh(**MESHES[name]), tpamr.BlockTreeMesh(**MESHES[name])
    k = drop_kappa(om, pm, centre, radius)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), precision=precision)
    x, y = pm.cell_centers()
    b = techref_noise(pm.ncells) + 0.25
    x0 = 0.1 * np.sin(3 * x) * np.cos(2 * y) + 2.0
    xg = x0.copy()
    sg = gs.solve(xg, b, rtol=1e-10, max_cycles=300)
    xo, so = OSolver(om, k).solve(b, x0=x0, rtol=1e-10, max_cycles=300)
    scale = np.sum(np.abs(b) * pm.volumes) / np.sum(pm.volumes)
    assert abs(sg.removed_rhs_mean - so["removed_rhs_mean"]) <= 1e-13 * scale
    assert math.isclose(sg.initial_linf, so["initial_linf"], rel_tol=1e-10)
    assert math.isclose(sg.initial_l2, so["initial_l2"], rel_tol=1e-10)
    assert math.isclose(sg.target_linf, so["target_linf"], rel_tol=1e-10)
    assert math.isclose(sg.residual_scale, so["residual_scale"], rel_tol=1e-10)
    assert sg.converged and so["converged"]
    assert sg.final_linf <= sg.target_linf
    if precision == "fp64":
        assert sg.cycles == so["cycles"]
        # the last iterate's residual (here ~3e-11 of the initial one at
        # 1000:1) carries the round-off of both codes' V-cycles
        assert math.isclose(sg.final_linf, so["final_linf"], rel_tol=1e-2)
        assert math.isclose(sg.final_l2, so["final_l2"], rel_tol=1e-2)
    gs.close()


@pytest.mark.parametrize("name,centre,radius", STATS_CASES)
def test_roundoff_floor_target_matches_oracle(name, centre, radius):
    """rtol = 1e-15 is below what FP64 reaches at 1000:1: both codes stop on
    the guarded floor 128 eps max(1, |A||x|/V, |b|/V) of their own final
    iterate, which agree to their last digits."""
    _print("case floor", name)
    om, pm = OMesh(**MESHES[name]), tpamr.BlockTreeMesh(**MESHES[name])
    k = drop_kappa(om, pm, centre, radius)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), precision="fp64")
    b = techref_noise(pm.ncells)
    xg = np.zeros(pm.ncells)
    sg = gs.solve(xg, b, rtol=1e-15, max_cycles=500)
    os_ = OSolver(om, k)
    xo, so = os_.solve(b, rtol=1e-15, max_cycles=500)
    assert sg.converged and so["converged"]
    rtol_target = 1e-15 * sg.initial_linf
    # the floor decided both targets
    assert sg.target_linf > rtol_target and so["target_linf"] > rtol_target
    assert math.isclose(sg.target_linf, so["target_linf"], rel_tol=1e-8)
    # the library's target is the oracle's floor of the library's iterate
    bint = b * pm.volumes
    bint -= pm.volumes * sg.removed_rhs_mean
    assert math.isclose(sg.target_linf, os_.floor(xg, bint), rel_tol=1e-12)
    assert sg.final_linf <= sg.target_linf
    gs.close()


def test_periodic_nullspace_convention_on_gpu():
    """test/operators/poisson.jl:921-933 on the library: constant rhs 3 on a
    periodic 8x8 mesh, initial pressure 7, coarse_size 4."""
    _print("case nullspace")
    pm = tpamr.BlockTreeMesh(block_size=(8, 8))
    for precision in ("fp64", "fp32_vcycle", "fp32_refine"):
        gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm), precision=precision,
                                  multigrid=tpamr.MultigridOptions(coarse_size=4))
        rhs = np.full(pm.ncells, 3.0)
        with pytest.raises(ValueError, match="nonzero volume mean"):
            gs.solve(np.full(pm.ncells, 7.0), rhs, compatibility="error")
        with pytest.raises(ValueError):
            gs.solve(np.full(pm.ncells, 7.0), rhs, compatibility="invalid")
        p = np.full(pm.ncells, 7.0)
        st = gs.solve(p, rhs, compatibility="project")
        eps = np.finfo(float).eps
        assert st.converged
        assert math.isclose(st.removed_rhs_mean, 3.0, rel_tol=math.sqrt(eps))
        assert abs(np.dot(pm.volumes, p)) <= 16 * eps
        assert np.max(np.abs(p)) <= 16 * eps
        gs.close()


def _shim_refined_keys(pm):
    """`_refined_keys` of the Julia shim: every proper ancestor of every leaf."""
    seen = set()
    for level, i, j in pm.block_keys:
        level, i, j = int(level), int(i), int(j)
        while level > 0:
            level, i, j = level - 1, i >> 1, j >> 1
            if (level, i, j) in seen:
                break
            seen.add((level, i, j))
    return np.asarray(sorted(seen), np.int32).reshape(-1, 3)


def test_shim_solve_override_mirror():
    """The calls of the Julia shim's `solve!` override (julia/TPAMRShim.jl),
    in order and with its arguments: tpamr_create (FP32 refinement),
    tpamr_set_multigrid, tpamr_set_mesh from the leaves' ancestors,
    tpamr_set_face_coefficients, tpamr_solve on the strong rhs and the
    initial pressure. The pressure matches the oracle's FP64 solve."""
    _print("case shim")
    kw = MESHES["amr64"]
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    k = drop_kappa(om, pm)
    lib = L.lib()
    h = C.c_void_p()
    L.check(lib.tpamr_create(C.byref(h), 0, L.TPAMR_FP32_REFINE), h)
    try:
        opts = L.MGOptions(64, 2, 2, 32)
        L.check(lib.tpamr_set_multigrid(h, C.byref(opts)), h)
        keys = np.ascontiguousarray(_shim_refined_keys(pm))
        base = np.asarray(pm.base_mesh, np.int32)
        bs = np.asarray(pm.block_size, np.int32)
        org = np.asarray(pm.origin, np.float64)
        ext = np.asarray(pm.extent, np.float64)
        per = np.asarray([int(p) for p in pm.periodic], np.int32)
        L.check(lib.tpamr_set_mesh(h, L.ptr(base), L.ptr(bs), L.ptr(org), L.ptr(ext), L.ptr(per), pm.max_level,
                                   len(keys), L.ptr(keys)), h)
        L.check(lib.tpamr_set_face_coefficients(h, L.ptr(np.ascontiguousarray(k))), h)
        b = techref_noise(pm.ncells)
        x = np.full(pm.ncells, 0.5)
        st = L.Stats()
        L.check(lib.tpamr_solve(h, L.ptr(b), L.ptr(x), 1e-10, 0.0, 100, 0, C.byref(st)), h)
        xo, so = OSolver(om, k).solve(b, x0=np.full(pm.ncells, 0.5), rtol=1e-10, max_cycles=100)
        assert st.converged and st.final_linf <= st.target_linf
        assert rel(x, xo) < 1e-7
        assert math.isclose(st.target_linf, so["target_linf"], rel_tol=1e-11)
    finally:
        lib.tpamr_destroy(h)


def test_coarsest_level_above_8192_cells():
    """A 3 x 3 layout of 32^2 roots cannot agglomerate: the hierarchy stops at
    9216 cells, solved by the dense inverse from the blocked Gauss-Jordan.
    The pressure equals the exact solution of the oracle's assembled operator
    (sparse direct solve of the augmented system [A V; V^T 0])."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    _print("case coarse 9216")
    kw = dict(base_mesh=(96, 96), block_size=(32, 32), periodic=(False, False), max_level=1,
              refine=lambda k, o, e: k[0] == 0 and (k[1] + k[2]) % 2 == 0)
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    k = np.exp(0.5 * np.sin(np.arange(1, pm.nfaces + 1)))
    A = OSolver(om, k, multigrid=False).level_matrix(0).tocsc()
    V = om.volumes
    n = pm.ncells
    aug = sp.bmat([[A, sp.csc_matrix(V.reshape(-1, 1))], [sp.csc_matrix(V.reshape(1, -1)), None]]).tocsc()
    b = techref_noise(n)
    bi = b * V
    bi -= V * bi.sum() / V.sum()
    exact = spla.spsolve(aug, np.concatenate([bi, [0.0]]))[:n]
    for precision in ("fp64", "fp32_refine"):
        gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), precision=precision)
        sizes = [lv["ncells"] for lv in gs.levels]
        assert sizes[-1] == 9216, sizes
        x = np.zeros(n)
        st = gs.solve(x, b, rtol=1e-12, max_cycles=100)
        assert st.converged
        assert rel(x, exact) < 1e-9, (precision, rel(x, exact))
        gs.close()


def test_coarsest_level_above_16384_cells():
    """A 5 x 5 layout of 32^2 roots stops at 25 600 cells (the reference
    factors any coarsest size, geometric_preconditioner.jl:6-24): a 25 601^2
    dense inverse (5.2 GB) from the blocked Gauss-Jordan. The pressure equals
    the exact solution of the oracle's assembled operator."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    _print("case coarse 25600")
    kw = dict(base_mesh=(160, 160), block_size=(32, 32), periodic=(False, False), max_level=1,
              refine=lambda k, o, e: k[0] == 0 and (k[1] + 2 * k[2]) % 3 == 0)
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    k = np.exp(0.5 * np.cos(np.arange(1, pm.nfaces + 1)))
    A = OSolver(om, k, multigrid=False).level_matrix(0).tocsc()
    V = om.volumes
    n = pm.ncells
    aug = sp.bmat([[A, sp.csc_matrix(V.reshape(-1, 1))], [sp.csc_matrix(V.reshape(1, -1)), None]]).tocsc()
    b = techref_noise(n)
    bi = b * V
    bi -= V * bi.sum() / V.sum()
    exact = spla.spsolve(aug, np.concatenate([bi, [0.0]]))[:n]
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), precision="fp32_refine")
    sizes = [lv["ncells"] for lv in gs.levels]
    assert sizes[-1] == 25600, sizes
    x = np.zeros(n)
    st = gs.solve(x, b, rtol=1e-12, max_cycles=100)
    assert st.converged
    assert rel(x, exact) < 1e-9, rel(x, exact)
    gs.close()
