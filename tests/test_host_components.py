"""The library's host hierarchy on domains split by zero face coefficients,
against the oracle (CPU, host-only context): every level's components (slots
and roots, each ungauged component's smallest cell), level operators, Jacobi
weights and smoothed transfers.

Meshes: the reference's component-label mesh (test/operators/poisson.jl:
1160-1244: 8 x 8 walled cells in 4 x 4 blocks, barriers of zero x-face
coefficients at x = 1/4, 3/4), a 64^2 AMR mesh with a zero-coefficient ring
that cuts out a disk (an inner component at every level it survives), and
the same with open left/right sides (the outer component gauged)."""
import numpy as np
import pytest

import tpamr
from oracle import OMesh, OSolver


def _strips(m):
    return np.where((m.face_axis == 0) & np.isin(m.face_center[:, 0], [0.25, 0.75]), 0.0, 1.0)


def _ring(m):
    """zero coefficient on the faces whose centre lies within 0.02 of the circle r = 0.3 about (0.5, 0.5)"""
    r = np.hypot(m.face_center[:, 0] - 0.5, m.face_center[:, 1] - 0.5)
    k = np.exp(0.3 * np.sin(np.arange(1, m.nfaces + 1)))
    return np.where(np.abs(r - 0.3) < 0.02, 0.0, k)


CASES = {
    "strips": (dict(base_mesh=(8, 8), block_size=(4, 4), periodic=(False, False)), _strips, None, 8),
    "ring_amr": (dict(base_mesh=(64, 64), block_size=(8, 8), periodic=(False, False), max_level=1,
                      refine=lambda k, o, e: o[0] < 0.5 and o[1] < 0.5), _ring, None, 64),
    "ring_amr_open": (dict(base_mesh=(64, 64), block_size=(8, 8), periodic=(False, False), max_level=1,
                           refine=lambda k, o, e: o[0] < 0.5 and o[1] < 0.5), _ring, (True, True, False, False), 64),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_components_and_levels_match_oracle(name):
    kw, kfun, sides, coarse = CASES[name]
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    k = kfun(om)
    assert np.any(k == 0)
    os_ = OSolver(om, k, coarse_size=coarse, open_sides=sides)
    ph = tpamr.HostHierarchy(pm, k, multigrid=tpamr.MultigridOptions(coarse_size=coarse), velocity_boundary=sides)
    assert [os_.level_info(l)["ncells"] for l in range(os_.nlevels)] == ph.sizes()
    several = False
    for l in range(os_.nlevels):
        labels, anchors = os_.components(l)
        slot, roots = ph.level_components(l)
        assert list(roots) == [int(a) + 1 for a in anchors], l
        # slot of every cell: its root's rank among the ungauged roots
        rank = {int(a): q + 1 for q, a in enumerate(anchors)}
        assert np.array_equal(slot, [rank.get(int(r), 0) for r in labels]), l
        several = several or len(anchors) > 1 or (len(anchors) == 1 and len(set(labels.tolist())) > 1)
        A, B = os_.level_matrix(l).tocsr(), ph.level_matrix(l)
        assert abs(A - B).max() <= 1e-14 * abs(A).max(), l
        assert abs(os_.level_info(l)["omega"] - ph.level_info(l)["omega"]) <= 1e-15, l
        if l + 1 < os_.nlevels:
            assert abs(os_.level_matrix(l, "P").tocsr() - ph.prolongation(l)).max() <= 1e-14, l
    assert several
    nl = ph.nlevels
    assert ph.level_boundary(nl - 1)["coarse_augmented"] == (len(os_.components(nl - 1)[1]) > 0)


def test_negative_coefficients_raise():
    m = tpamr.BlockTreeMesh(block_size=(8, 8))
    k = np.ones(m.nfaces)
    k[3] = -1.0
    with pytest.raises(ValueError, match="nonnegative"):
        tpamr.HostHierarchy(m, k)
