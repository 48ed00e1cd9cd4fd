"""CPU pins of the oracle's pressure components (no GPU): connected components
over faces with a nonzero coefficient, each ungauged one projected and
gauged on its own.

Restated TwoPhaseAMR.jl test:
  open-face Poisson component labels      test/operators/poisson.jl:1160-1244
(an 8 x 8 walled mesh in 4 x 4 blocks split by zero-coefficient x faces into
strips: labels are each strip's smallest cell, the strip-constant rhs +-1 is
all nullspace, so the removed component means are exactly those constants,
the pressure is zero, and a refresh to unit coefficients merges the strips).
Beside it: with open sides the strips that touch an open face are gauged and
keep no slot; the multigrid preconditioner gives the plain-CG answers.
"""
import numpy as np
import pytest

from oracle import OMesh, OSolver


def _mesh():
    return OMesh(base_mesh=(8, 8), block_size=(4, 4), periodic=(False, False))


def _barriers(m, xs):
    return np.where((m.face_axis == 0) & np.isin(m.face_center[:, 0], xs), 0.0, 1.0)


def _strips(m, cuts):
    x, _ = m.cell_centers()
    edges = [-np.inf, *cuts, np.inf]
    return [np.flatnonzero((x > lo) & (x < hi)) for lo, hi in zip(edges[:-1], edges[1:])]


@pytest.mark.parametrize("multigrid", [False, True])
def test_two_strips_labels(multigrid):
    m = _mesh()
    s = OSolver(m, _barriers(m, [0.5]), coarse_size=8, multigrid=multigrid)
    labels, anchors = s.components(0)
    left, right = _strips(m, [0.5])
    expected = np.zeros(m.ncells, np.int64)
    expected[left] = left.min()
    expected[right] = right.min()
    assert np.array_equal(labels, expected)
    assert list(anchors) == sorted([left.min(), right.min()])
    labels, anchors = OSolver(m, np.ones(m.nfaces), coarse_size=8, multigrid=multigrid).components(0)
    assert np.all(labels == 0) and list(anchors) == [0]


@pytest.mark.parametrize("multigrid", [False, True])
def test_three_strips_projection(multigrid):
    m = _mesh()
    k = _barriers(m, [0.25, 0.75])
    strips = _strips(m, [0.25, 0.75])
    x, _ = m.cell_centers()
    rhs = np.where((x < 0.25) | (x > 0.75), 1.0, -1.0)
    s = OSolver(m, k, coarse_size=8, multigrid=multigrid)
    labels, anchors = s.components(0)
    assert list(anchors) == sorted(c.min() for c in strips)
    with pytest.raises(RuntimeError):
        s.solve(rhs, compatibility="error")
    p, st = s.solve(rhs, x0=np.full(m.ncells, 7.0), compatibility="project", rtol=0.0, atol=1e-12)
    assert st["converged"]
    assert st["removed_rhs_mean"] == 0.0
    assert st["removed_rhs_component_anchors"] == [int(c.min()) + 1 for c in strips]
    assert st["removed_rhs_component_means"] == [1.0, -1.0, 1.0]
    assert np.all(p == 0)
    # merged (unit coefficients): one component, nothing to remove
    merged = OSolver(m, np.ones(m.nfaces), coarse_size=8, multigrid=multigrid)
    p, st = merged.solve(rhs, compatibility="error", rtol=0.0, atol=1e-10, max_cycles=500)
    assert st["converged"] and st["removed_rhs_component_anchors"] == [1]
    assert abs(st["removed_rhs_component_means"][0]) <= 1e-12


def test_strips_with_a_generic_rhs_match_per_strip_solves():
    """Each strip's solution is the solution of its own Neumann problem with
    its own mean removed: a solve on the split mesh equals, strip by strip,
    the per-strip projections of the rhs and zero-mean pressures."""
    m = _mesh()
    k = _barriers(m, [0.25, 0.75])
    strips = _strips(m, [0.25, 0.75])
    x, y = m.cell_centers()
    rhs = np.sin(3 * x + 1) * np.cos(2 * y) + 0.2
    s = OSolver(m, k, coarse_size=8)
    p, st = s.solve(rhs, rtol=1e-12, max_cycles=200)
    assert st["converged"]
    V = m.volumes
    A = s.level_matrix(0).tocsr()
    for c, mean in zip(strips, st["removed_rhs_component_means"]):
        assert mean == pytest.approx(np.sum(rhs[c] * V[c]) / np.sum(V[c]), rel=1e-13)
        assert abs(np.dot(V[c], p[c])) <= 1e-13 * np.sum(V[c])
        b = (rhs[c] - mean) * V[c]
        r = b - (A @ p)[c]
        assert np.max(np.abs(r / V[c])) <= 1e-10 * np.max(np.abs(b / V[c]))


def test_open_side_gauges_its_strips():
    """Open left and right sides: the outer strips touch an open face and are
    gauged; only the middle strip is projected."""
    m = _mesh()
    k = _barriers(m, [0.25, 0.75])
    strips = _strips(m, [0.25, 0.75])
    x, y = m.cell_centers()
    rhs = np.sin(3 * x + 1) * np.cos(2 * y) + 0.2
    s = OSolver(m, k, coarse_size=8, open_sides=(True, True, False, False))
    labels, anchors = s.components(0)
    assert list(anchors) == [strips[1].min()]
    p, st = s.solve(rhs, rtol=1e-12, max_cycles=200)
    assert st["converged"]
    V = m.volumes
    mid = strips[1]
    assert st["removed_rhs_component_anchors"] == [int(mid.min()) + 1]
    assert st["removed_rhs_component_means"][0] == pytest.approx(np.sum(rhs[mid] * V[mid]) / np.sum(V[mid]),
                                                                 rel=1e-13)
    assert st["removed_rhs_mean"] == pytest.approx(st["removed_rhs_component_means"][0], rel=1e-15)
    A = s.level_matrix(0).tocsr()
    b = rhs * V
    b[mid] -= st["removed_rhs_component_means"][0] * V[mid]
    assert np.max(np.abs((b - A @ p) / V)) <= 1e-10 * np.max(np.abs(b / V))
