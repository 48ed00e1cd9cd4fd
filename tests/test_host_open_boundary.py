"""The library's host hierarchy with open (homogeneous Dirichlet) pressure
sides against the oracle (CPU, host-only context): every level's operator
(the open faces' boundary diagonal in regular slots and in irregular rows),
Jacobi weights, smoothed transfers, boundary coefficients restricted to the
coarse levels, singularity of every level and the coarse solve's volume
constraint.

Meshes: the reference's open-boundary test mesh (test/operators/poisson.jl:
935-950; 16 x 8 cells in 4 x 4 blocks on 2 x 1, one refined root column
beside the outflow side, periodic or not in y), the same with a refined
corner root (coarse-fine interfaces meeting the open sides, so irregular rows
carry open faces), and a 64^2 AMR case with all four sides open."""
import numpy as np
import pytest

import tpamr
from oracle import OMesh, OSolver

from cases import OPEN_CASES as CASES, OPEN_LR


def _kappa(m):
    return np.exp(0.3 * np.sin(np.arange(1, m.nfaces + 1)))


@pytest.mark.parametrize("name", sorted(CASES))
def test_open_hierarchy_matches_oracle(name):
    kw, sides, bkind = CASES[name]
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    assert np.array_equal(om.bface_cell, pm.bface_cell) and np.array_equal(om.bface_axis, pm.bface_axis)
    k = _kappa(pm)
    kb = None if bkind == "extended" else 0.8 + 0.2 * om.bface_center[:, 1] + 0.1 * om.bface_center[:, 0]
    os_ = OSolver(om, k, coarse_size=8, open_sides=sides, boundary_coefficient=kb)
    ph = tpamr.HostHierarchy(pm, k, multigrid=tpamr.MultigridOptions(coarse_size=8),
                             velocity_boundary=sides, boundary_face_coefficient=kb)
    assert [os_.level_info(l)["ncells"] for l in range(os_.nlevels)] == ph.sizes()
    for l in range(os_.nlevels):
        ob, hb = os_.level_boundary(l), ph.level_boundary(l)
        assert np.allclose(hb["coefficient"], ob["coefficient"], rtol=1e-14, atol=0), l
        assert hb["singular"] == ob["singular"] is False, l
        A, B = os_.level_matrix(l).tocsr(), ph.level_matrix(l)
        assert abs(A - B).max() <= 1e-14 * abs(A).max(), l
        # the boundary diagonal is the row sum of the anchored operator
        assert np.allclose(np.asarray(B.sum(axis=1)).ravel(), ob["diagonal"], rtol=1e-10, atol=1e-12 * abs(A).max())
        assert abs(os_.level_info(l)["omega"] - ph.level_info(l)["omega"]) <= 1e-15, l
        assert np.allclose(os_.level_kappa(l), ph.level_kappa(l), rtol=1e-14, atol=0)
        if l + 1 < os_.nlevels:
            assert abs(os_.level_matrix(l, "P").tocsr() - ph.prolongation(l)).max() <= 1e-14
    assert not ph.level_boundary(ph.nlevels - 1)["coarse_augmented"]


def test_zero_boundary_coefficients_stay_singular():
    """Open sides whose coefficients are all zero anchor nothing: the
    operator keeps its nullspace and the coarse solve its constraint
    (elliptic_solver.jl:60-109)."""
    kw, sides, _ = CASES["outflow_column"]
    pm = tpamr.BlockTreeMesh(**kw)
    ph = tpamr.HostHierarchy(pm, 1.0, multigrid=tpamr.MultigridOptions(coarse_size=8), velocity_boundary=sides,
                             boundary_face_coefficient=0.0)
    for l in range(ph.nlevels):
        assert ph.level_boundary(l)["singular"]
    assert ph.level_boundary(ph.nlevels - 1)["coarse_augmented"]
    A = ph.level_matrix(0)
    assert abs(np.asarray(A.sum(axis=1)).ravel()).max() <= 1e-12 * abs(A).max()


def test_open_side_errors():
    kw, sides, _ = CASES["outflow_column"]
    pm = tpamr.BlockTreeMesh(**kw)
    with pytest.raises(ValueError):
        tpamr.HostHierarchy(pm, 1.0, velocity_boundary=sides, boundary_face_coefficient=-1.0)
    with pytest.raises(ValueError):
        tpamr.HostHierarchy(pm, 1.0, velocity_boundary=sides, boundary_face_coefficient=np.nan)
    with pytest.raises(ValueError):
        tpamr.pressure_boundary_sides((True, False))
    vb = tpamr.VelocityBoundaryConditions(left=tpamr.ZeroGradientVelocity(), right=tpamr.Outflow(),
                                          bottom=tpamr.NoSlip(), top=tpamr.FreeSlip())
    assert tpamr.pressure_boundary_sides(vb) == OPEN_LR
    assert tpamr.pressure_boundary_sides(None) == (False,) * 4


def test_mixture_boundary_coefficients_match_oracle():
    """`_mixture_boundary_coefficient` (forces.jl:103-111): 1 / (rho_l c + rho_g
    (1 - c)) of each boundary face's cell, bit for bit with the oracle (and
    the closed form)."""
    import ctypes as C
    from tpamr import _lib as L
    kw, _, _ = CASES["amr64_all_open"]
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    x, y = pm.cell_centers()
    c = np.clip(0.5 + 2.0 * (0.3 - np.hypot(x - 0.5, y - 0.5)), 0.0, 1.0)
    kb = np.zeros(pm.nboundary_faces)
    L.check(L.lib().tpamr_mixture_boundary_coefficients(pm._host.h, L.ptr(c), 1000.0, 1.2, L.ptr(kb)), pm._host.h)
    assert np.array_equal(kb, om.mixture_boundary_coefficients(c, 1000.0, 1.2))
    cb = c[pm.bface_cell]
    assert np.array_equal(kb, 1.0 / (1000.0 * cb + 1.2 * (1.0 - cb)))
