"""GPU: open (homogeneous Dirichlet) pressure sides through the C ABI against
the CPU FP64 oracle, whose open-side operator and solver are pinned to the
reference's own tests in tests/test_oracle_open_boundary.py.

- The device operator (regular slots and irregular rows with open faces) and
  the V-cycle-preconditioned solve in every precision mode, on the meshes of
  tests/cases.py OPEN_CASES: the manufactured pressure of
  test/operators/poisson.jl:994-1009 is recovered from its nonzero-mean
  strong rhs with compatibility = :error, no mean is removed, FP64 takes the
  oracle's iteration count.
- Coefficient refresh on the device keeps the boundary diagonal
  (poisson.jl:1011-1035): new interior and boundary coefficients, every level
  against a fresh oracle.
- Boundary gradient and the projection of boundary velocities; the
  interior-face projection rejects an operator with open sides and the open
  one an operator without (poisson.jl:1037-1045).
- One context switching from closed to open sides and back.
- At size: the 64-circle mixture mesh (levels 5-8, 3 659 008 cells, density
  ratio 1000) with outflow on the left and right sides, against the oracle.
"""
import numpy as np
import pytest

from cases import OPEN_CASES, OPEN_LR

pytestmark = pytest.mark.gpu

import tpamr  # noqa: E402
from oracle import OMesh, OSolver  # noqa: E402

MODES = ("fp64", "fp32_vcycle", "fp32_refine")


def _print(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _kappa(m):
    return np.exp(0.3 * np.sin(np.arange(1, m.nfaces + 1)))


def _setup(name, coarse_size=8):
    kw, sides, bkind = OPEN_CASES[name]
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    k = _kappa(pm)
    kb = None if bkind == "extended" else 0.8 + 0.2 * om.bface_center[:, 1] + 0.1 * om.bface_center[:, 0]
    os_ = OSolver(om, k, coarse_size=coarse_size, open_sides=sides, boundary_coefficient=kb)
    return om, pm, k, kb, sides, os_


def _pressure(pm):
    x, y = pm.cell_centers()
    return 0.7 + np.sin(np.pi * x / 2) * np.cos(np.pi * y)


@pytest.mark.parametrize("name", sorted(OPEN_CASES))
def test_open_operator_and_solve(name):
    _print("case open", name)
    om, pm, k, kb, sides, os_ = _setup(name)
    p = _pressure(pm)
    Ap = os_.apply(p)
    strong = Ap / pm.volumes
    xo, so = os_.solve(strong, compatibility="error", rtol=1e-11, max_cycles=300)
    assert so["converged"] and so["removed_rhs_mean"] == 0
    for prec in MODES:
        s = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k, velocity_boundary=sides, boundary_face_coefficient=kb),
                                 precision=prec, coarse_size=8)
        if prec == "fp64":
            assert rel(s.apply(p), Ap) < 1e-13
            gb = tpamr.weighted_gradient_boundary(s, p)
            assert np.allclose(gb, os_.weighted_gradient_boundary(p), rtol=1e-15, atol=0)
        assert not s.level_boundary(0)["singular"]
        assert not s.level_boundary(len(s.levels) - 1)["coarse_augmented"]
        x = np.zeros(pm.ncells)
        st = s.solve(x, strong, compatibility="error", rtol=1e-11, max_cycles=300)
        _print("  ", prec, st.cycles, so["cycles"], rel(x, xo), rel(x, p))
        assert st.converged and st.removed_rhs_mean == 0
        assert np.allclose(x, p, atol=1e-9, rtol=1e-9)
        if prec == "fp64":
            assert st.cycles == so["cycles"]
            assert rel(x, xo) < 1e-10
        else:
            assert rel(x, xo) < 1e-8
        s.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_open_device_refresh(precision):
    """The second coefficient update runs on the device: new interior (1.7
    times the first) and boundary (0.6) coefficients on every level, against
    a fresh oracle."""
    _print("case open refresh", precision)
    om, pm, k, kb, sides, _ = _setup("corner_refined")
    op = tpamr.PoissonOperator(pm, k, velocity_boundary=sides, boundary_face_coefficient=kb)
    s = tpamr.EllipticSolver(op, precision=precision, coarse_size=8)
    p = _pressure(pm)
    x = np.zeros(pm.ncells)
    s.solve(x, s.apply(p) / pm.volumes, rtol=1e-11, max_cycles=300)
    k2 = 1.7 * k
    op.replace_face_coefficient(k2, boundary_face_coefficient=np.full(pm.nboundary_faces, 0.6))
    os2 = OSolver(om, k2, coarse_size=8, open_sides=sides, boundary_coefficient=np.full(om.nbfaces, 0.6))
    Ap = os2.apply(p)
    if precision == "fp64":
        assert rel(s.apply(p), Ap) < 1e-13
    for l in range(os2.nlevels):
        assert np.allclose(s.level_boundary(l)["coefficient"], os2.level_boundary(l)["coefficient"], rtol=1e-14)
        assert abs(s.level_info(l)["omega"] - os2.level_info(l)["omega"]) <= 1e-15
    x = np.zeros(pm.ncells)
    st = s.solve(x, Ap / pm.volumes, compatibility="error", rtol=1e-11, max_cycles=300)
    assert st.converged and st.removed_rhs_mean == 0
    assert np.allclose(x, p, atol=1e-9, rtol=1e-9)
    # all-zero boundary coefficients: singular again, the rhs is projected
    op.replace_face_coefficient(k2, boundary_face_coefficient=np.zeros(pm.nboundary_faces))
    os3 = OSolver(om, k2, coarse_size=8, open_sides=sides, boundary_coefficient=np.zeros(om.nbfaces))
    rhs = np.cos(3 * pm.cell_centers()[0])
    x = np.zeros(pm.ncells)
    st = s.solve(x, rhs, rtol=1e-10, max_cycles=300)
    xo, so = os3.solve(rhs, rtol=1e-10, max_cycles=300)
    assert s.level_boundary(0)["singular"] and s.level_boundary(len(s.levels) - 1)["coarse_augmented"]
    assert st.converged and st.removed_rhs_mean == pytest.approx(so["removed_rhs_mean"], rel=1e-12)
    assert rel(x, xo) < (1e-10 if precision == "fp64" else 1e-8)
    s.close()


def test_open_projection_and_side_checks():
    _print("case open projection")
    om, pm, k, kb, sides, os_ = _setup("outflow_column_periodic_y")
    s = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k, velocity_boundary=sides, boundary_face_coefficient=kb),
                             coarse_size=8)
    rng = np.random.default_rng(11)
    u = rng.standard_normal(pm.nfaces)
    ub = np.where(pm.bface_axis == 0, rng.standard_normal(pm.nboundary_faces), 0.0)
    uo, ubo, po, phio, so = os_.project_open(u, ub, np.zeros(pm.ncells), 0.1, tolerance=1e-12, max_cycles=200)
    u2, ub2, p2 = u.copy(), ub.copy(), np.zeros(pm.ncells)
    st, phi = tpamr.project_open(u2, ub2, p2, s, 0.1, tolerance=1e-12, max_cycles=200)
    assert st.converged and st.removed_rhs_mean == 0
    assert rel(phi, phio) < 1e-10
    assert np.allclose(u2, uo, rtol=0, atol=1e-10 * np.max(np.abs(uo)))
    assert np.allclose(ub2, ubo, rtol=0, atol=1e-10 * np.max(np.abs(ubo)))
    assert np.max(np.abs(tpamr.divergence(s, u2, ub2))) <= 1e-8 * np.max(np.abs(tpamr.divergence(s, u, ub)))
    # interior-face projection of an operator with open sides, and the open
    # projection of one without: ArgumentError, inputs untouched
    before = u.copy()
    with pytest.raises(ValueError, match="open sides"):
        tpamr.project(u, np.zeros(pm.ncells), s, 0.1)
    assert np.array_equal(u, before)
    closed = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), coarse_size=8)
    with pytest.raises(ValueError, match="open sides"):
        tpamr.project_open(u, ub.copy(), np.zeros(pm.ncells), closed, 0.1)
    assert np.array_equal(u, before)
    closed.close()
    s.close()


def test_one_context_closed_open_closed():
    """Open sides set and cleared on one solver: each change rebuilds the
    hierarchy, each solve matches the oracle of its sides."""
    _print("case open toggle")
    kw, sides, _ = OPEN_CASES["amr64_all_open"]
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    k = _kappa(pm)
    rhs = np.sin(2 * np.pi * pm.cell_centers()[0]) + 0.3
    op = tpamr.PoissonOperator(pm, k)
    s = tpamr.EllipticSolver(op, precision="fp64", coarse_size=8)
    for open_sides in (None, sides, None):
        if open_sides is None:
            op.boundary_sides = (False,) * 4
            op.boundary_face_coefficient = None
        else:
            op.boundary_sides = open_sides
            op.boundary_face_coefficient = None  # extended from the interior coefficients
        op.version += 1
        os_ = OSolver(om, k, coarse_size=8, open_sides=open_sides)
        x = np.zeros(pm.ncells)
        st = s.solve(x, rhs, rtol=1e-10, max_cycles=300)
        xo, so = os_.solve(rhs, rtol=1e-10, max_cycles=300)
        _print("  sides", open_sides, st.cycles, so["cycles"], rel(x, xo))
        assert st.converged and st.cycles == so["cycles"]
        assert rel(x, xo) < 1e-10
        assert st.removed_rhs_mean == pytest.approx(so["removed_rhs_mean"], rel=1e-12, abs=0)
    s.close()


def test_open_circles_levels5_8_against_oracle():
    """At size: 3 659 008 cells, density ratio 1000, outflow left and right
    (boundary coefficients the mixture's inverse density of the boundary
    cells' fractions), every mode against the oracle's FP64 solution."""
    import torch
    from tpamr import workloads as W
    _print("case open circles 5-8: setup")
    mesh, circles = W.circles_mesh((5, 8))
    frac = W.liquid_fraction(mesh, circles)
    om = OMesh(base_mesh=mesh.base_mesh, block_size=mesh.block_size, origin=mesh.origin, extent=mesh.extent,
               periodic=mesh.periodic, max_level=mesh.max_level, refined_keys=mesh.refined_keys)
    kappa = om.mixture_coefficients(frac, 1000.0, 1.0)
    kb = om.mixture_boundary_coefficients(frac, 1000.0, 1.0)
    sides = (True, True, False, False)
    u = W.smooth_velocity(mesh, device="cuda:0")
    s0 = tpamr.EllipticSolver(tpamr.PoissonOperator(mesh, kappa), precision="fp64")
    rhs = -tpamr.divergence(s0, u) + 0.1
    s0.close()
    xo, so = OSolver(om, kappa, open_sides=sides, boundary_coefficient=kb).solve(rhs, rtol=1e-10, max_cycles=200)
    assert so["converged"]
    for prec in MODES:
        # the two-phase refresh sets the mixture's interior and boundary coefficients
        s = tpamr.EllipticSolver(tpamr.PoissonOperator(mesh, 1.0, velocity_boundary=sides), precision=prec)
        tpamr.refresh_two_phase(s, frac, (1000.0, 1.0))
        assert np.array_equal(s.operator.boundary_face_coefficient, kb)
        assert np.array_equal(s.operator.face_coefficient, kappa)
        d_rhs = torch.as_tensor(rhs, device="cuda:0")
        d_x = torch.zeros(mesh.ncells, dtype=torch.float64, device="cuda:0")
        st = s.solve_device(d_rhs.data_ptr(), d_x.data_ptr(), rtol=1e-10, max_cycles=200)
        torch.cuda.synchronize()
        x = d_x.cpu().numpy()
        _print("  ", prec, st.cycles, so["cycles"], rel(x, xo), st.solve_ms)
        assert st.converged and st.removed_rhs_mean == 0
        assert rel(x, xo) < 1e-9, prec
        if prec == "fp64":
            assert st.cycles == so["cycles"]
        s.close()
        del d_x, d_rhs
        torch.cuda.empty_cache()
