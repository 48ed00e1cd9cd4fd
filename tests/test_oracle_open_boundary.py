"""CPU pins of the oracle's open (homogeneous Dirichlet) pressure sides
(no GPU).

Restated TwoPhaseAMR.jl tests:
  open pressure boundary operators and solvers            test/operators/poisson.jl:935-1046
  vertical open faces with a periodic transverse axis     test/operators/poisson.jl:1052-1074
Closed forms beside them: the boundary diagonal is the sum over a cell's open
faces of |f| 2 k_b / h; a coarse boundary coefficient of the linear material
0.8 + 0.2 y is that material at the coarse face centre (the restriction
averages the fine faces it covers); a constant coefficient extends to itself.
"""
import numpy as np
import pytest

from oracle import OMesh, OSolver

# left = ZeroGradientVelocity, right = Outflow, bottom = NoSlip, top = FreeSlip
OPEN_LR = (True, True, False, False)


def _mesh(mixed, periodic):
    return OMesh(base_mesh=(16, 8), block_size=(4, 4), extent=(2.0, 1.0), periodic=periodic, max_level=1,
                 refine=lambda k, o, e: mixed and k[0] == 0 and k[1] == 3)


def _spacing(m, cells, axis):
    nb = m.cells_per_block
    return m.block_geom[np.asarray(cells) // nb, 4 + axis]


def _pressure(m):
    x, y = m.cell_centers()
    return 0.7 + np.sin(np.pi * x / 2) * np.cos(np.pi * y)


CASES = [(mixed, per) for mixed in (False, True) for per in ((False, False), (False, True))]


@pytest.mark.parametrize("mixed,periodic", CASES)
def test_open_boundary_operator_gradient_and_solve(mixed, periodic):
    m = _mesh(mixed, periodic)
    material = 0.8 + 0.2 * m.bface_center[:, 1]
    s = OSolver(m, 1.2, coarse_size=8, open_sides=OPEN_LR, boundary_coefficient=material)
    lb = s.level_boundary(0)
    assert not lb["singular"]
    # boundary diagonal: |f| 2 k_b / h over the open faces (x faces here)
    expect = np.zeros(m.ncells)
    h = _spacing(m, m.bface_cell, 0)
    openf = m.bface_axis == 0
    np.add.at(expect, m.bface_cell[openf], m.bface_area[openf] * 2 * material[openf] / h[openf])
    assert np.allclose(lb["diagonal"], expect, rtol=1e-15, atol=0)
    # factorized action == assembled matrix (poisson.jl:958-961)
    p = _pressure(m)
    A = s.level_matrix(0)
    assert np.allclose(s.apply(p), A @ p, atol=1e-12, rtol=1e-12)
    # interior coefficients extended to the open faces (poisson.jl:966-974)
    ext = OSolver(m, np.full(m.nfaces, 1.2), coarse_size=8, open_sides=OPEN_LR).level_boundary(0)["coefficient"]
    assert np.allclose(ext[openf], 1.2, rtol=1e-15)
    assert np.all(ext[~openf] == 0)
    # weighted gradient: -div(G p) is the strong operator; open faces carry
    # -side k_b p / (h / 2), closed faces zero (poisson.jl:976-992)
    g = s.weighted_gradient(p)
    gb = s.weighted_gradient_boundary(p)
    strong = s.apply(p) / m.volumes
    assert np.allclose(strong, -m.divergence(g, gb), atol=2e-12, rtol=0)
    expected_gb = np.where(openf, -m.bface_side * material * p[m.bface_cell] / (h / 2), 0.0)
    assert np.allclose(gb, expected_gb, atol=1e-12, rtol=0)
    assert np.all(gb[~openf] == 0)
    # a nonzero-mean rhs needs no compatibility projection (poisson.jl:994-1009)
    x, st = s.solve(strong, compatibility="error", rtol=1e-11, max_cycles=300)
    assert st["converged"]
    assert st["removed_rhs_mean"] == 0
    assert np.allclose(x, p, atol=1e-9, rtol=1e-9)
    assert not s.level_boundary(s.nlevels - 1)["coarse_augmented"]
    for l in range(s.nlevels):
        lev = s.level_boundary(l)
        assert not lev["singular"], l
        # coarse boundary coefficients: the linear material at the coarse
        # face centres
        lm = s.level_mesh(l)
        lo = lm.bface_axis == 0
        assert np.allclose(lev["coefficient"][lo], 0.8 + 0.2 * lm.bface_center[lo, 1], rtol=1e-14, atol=0)


@pytest.mark.parametrize("mixed,periodic", CASES)
def test_open_boundary_refreshed_coefficients(mixed, periodic):
    """poisson.jl:1011-1035: new interior (1.7) and boundary (0.6)
    coefficients keep the boundary diagonal on every level; the solve
    recovers the manufactured pressure."""
    m = _mesh(mixed, periodic)
    s = OSolver(m, 1.7, coarse_size=8, open_sides=OPEN_LR, boundary_coefficient=np.full(m.nbfaces, 0.6))
    p = _pressure(m)
    assert np.allclose(s.apply(p), s.level_matrix(0) @ p, atol=1e-12, rtol=1e-12)
    for l in range(s.nlevels):
        lev = s.level_boundary(l)
        lm = s.level_mesh(l)
        assert not lev["singular"]
        assert np.allclose(lev["coefficient"][lm.bface_axis == 0], 0.6, rtol=1e-15)
    x, st = s.solve(s.apply(p) / m.volumes, compatibility="error", rtol=1e-11, max_cycles=300)
    assert st["converged"]
    assert np.allclose(x, p, atol=1e-9, rtol=1e-9)


def test_vertical_open_faces_periodic_transverse():
    """poisson.jl:1052-1074: bottom ZeroGradientVelocity and top Outflow on a
    periodic-in-x 8 x 4 block, coefficient 2, constant pressure 1."""
    m = OMesh(block_size=(8, 4), extent=(3.0, 1.0), periodic=(True, False))
    s = OSolver(m, 2.0, open_sides=(False, False, True, True))
    p = np.ones(m.ncells)
    g = s.weighted_gradient(p)
    gb = s.weighted_gradient_boundary(p)
    assert np.all(g == 0)
    assert np.all(m.bface_axis == 1)
    h = _spacing(m, m.bface_cell, 1)
    assert np.allclose(gb, -m.bface_side * 2 / (h / 2), rtol=1e-15)
    strong = s.apply(p) / m.volumes
    assert np.allclose(strong, -m.divergence(g, gb), atol=1e-12, rtol=0)


def test_open_projection_corrects_boundary_velocities():
    """_project_prepared! (navier_stokes/projection.jl:221-260) with open
    sides: the boundary velocities enter the divergence and receive
    -dt G_b phi, so the projected field is discretely divergence free with
    its open-face fluxes."""
    m = _mesh(True, (False, True))
    s = OSolver(m, 1.2, coarse_size=8, open_sides=OPEN_LR)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(m.nfaces)
    ub = np.where(m.bface_axis == 0, rng.standard_normal(m.nbfaces), 0.0)
    u2, ub2, p2, phi, st = s.project_open(u, ub, np.zeros(m.ncells), 0.1, tolerance=1e-12, max_cycles=200)
    assert st["converged"] and st["removed_rhs_mean"] == 0
    div0 = np.max(np.abs(m.divergence(u, ub)))
    assert np.max(np.abs(m.divergence(u2, ub2))) <= 1e-9 * div0
    assert np.all(ub2[m.bface_axis == 1] == 0)
    assert np.array_equal(p2, phi)
