"""GPU: pressure components (domains split by zero face coefficients) through
the C ABI against the CPU FP64 oracle, whose components are pinned to the
reference's tests in tests/test_oracle_components.py.

- test/operators/poisson.jl:1160-1244 on the GPU in every precision mode: the
  strip-constant rhs is all nullspace (removed component means exactly
  1, -1, 1 with anchors at each strip's smallest cell, pressure zero from an
  initial 7, :error raises), a refresh to unit coefficients merges the strips
  (device refresh) and a refresh back splits them (host refresh).
- Generic right-hand sides on the strips and on a 64^2 AMR mesh with a
  zero-coefficient ring (an inner component on several levels, isolated cells
  with zero diagonals), walled and with open sides (the outer component
  gauged): pressures, iteration counts and removed component means against
  the oracle's literal V-cycle.
- The projection on a split domain: compatibility from the largest removed
  component mean and each ungauged component's own pressure gauge
  (_accumulate_pressure_increment!, projection.jl:52-97).
"""
import numpy as np
import pytest

from test_host_components import CASES as SPLIT_CASES

pytestmark = pytest.mark.gpu

import tpamr  # noqa: E402
from oracle import OMesh, OSolver  # noqa: E402

MODES = ("fp64", "fp32_vcycle", "fp32_refine")


def _print(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _strip_mesh():
    kw = dict(base_mesh=(8, 8), block_size=(4, 4), periodic=(False, False))
    return OMesh(**kw), tpamr.BlockTreeMesh(**kw)


@pytest.mark.parametrize("precision", MODES)
def test_strip_components_reference_case(precision):
    _print("case strips", precision)
    om, pm = _strip_mesh()
    k = np.where((om.face_axis == 0) & np.isin(om.face_center[:, 0], [0.25, 0.75]), 0.0, 1.0)
    x, _ = pm.cell_centers()
    strips = [np.flatnonzero(x < 0.25), np.flatnonzero((x > 0.25) & (x < 0.75)), np.flatnonzero(x > 0.75)]
    anchors = [int(c.min()) + 1 for c in strips]
    rhs = np.where((x < 0.25) | (x > 0.75), 1.0, -1.0)
    op = tpamr.PoissonOperator(pm, k)
    s = tpamr.EllipticSolver(op, precision=precision, coarse_size=8)
    with pytest.raises(ValueError, match="component"):
        s.solve(np.zeros(pm.ncells), rhs, compatibility="error")
    p = np.full(pm.ncells, 7.0)
    st = s.solve(p, rhs, compatibility="project", rtol=0.0, atol=1e-12)
    assert st.converged
    assert st.removed_rhs_mean == 0.0
    assert st.removed_rhs_component_anchors == anchors
    assert st.removed_rhs_component_means == [1.0, -1.0, 1.0]
    assert np.all(p == 0)
    # merged (a device refresh: every coefficient positive)
    op.replace_face_coefficient(np.ones(pm.nfaces))
    merged = s.solve(np.zeros(pm.ncells), rhs, compatibility="error", rtol=0.0, atol=1e-10, max_cycles=500)
    assert merged.converged and merged.removed_rhs_component_anchors == [1]
    assert abs(merged.removed_rhs_component_means[0]) <= 1e-12
    # split again (a host refresh)
    op.replace_face_coefficient(k)
    p = np.full(pm.ncells, 7.0)
    split = s.solve(p, rhs, rtol=0.0, atol=1e-12)
    assert split.converged and split.removed_rhs_component_anchors == anchors
    assert split.removed_rhs_component_means == [1.0, -1.0, 1.0]
    assert np.all(p == 0)
    s.close()


@pytest.mark.parametrize("name", sorted(SPLIT_CASES))
def test_split_domains_against_oracle(name):
    _print("case split", name)
    kw, kfun, sides, coarse = SPLIT_CASES[name]
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    k = kfun(om)
    x, y = pm.cell_centers()
    rhs = np.sin(3 * x + 1) * np.cos(2 * y) + 0.2
    os_ = OSolver(om, k, coarse_size=coarse, open_sides=sides)
    xo, so = os_.solve(rhs, rtol=1e-10, max_cycles=300)
    assert so["converged"]
    for prec in MODES:
        s = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k, velocity_boundary=sides), precision=prec,
                                 coarse_size=coarse)
        xg = np.zeros(pm.ncells)
        st = s.solve(xg, rhs, rtol=1e-10, max_cycles=300)
        _print("  ", prec, st.cycles, so["cycles"], rel(xg, xo))
        assert st.converged
        assert st.removed_rhs_component_anchors == so["removed_rhs_component_anchors"]
        assert np.allclose(st.removed_rhs_component_means, so["removed_rhs_component_means"], rtol=1e-12, atol=1e-15)
        assert st.removed_rhs_mean == pytest.approx(so["removed_rhs_mean"], rel=1e-12, abs=1e-15)
        if prec == "fp64":
            assert st.cycles == so["cycles"]
            assert rel(xg, xo) < 1e-10
        else:
            assert rel(xg, xo) < 1e-8
        # the device solve entry gives the same result
        if prec == "fp64":
            import torch
            d_rhs = torch.as_tensor(rhs, device="cuda:0")
            d_x = torch.zeros(pm.ncells, dtype=torch.float64, device="cuda:0")
            s.solve_device(d_rhs.data_ptr(), d_x.data_ptr(), rtol=1e-10, max_cycles=300)
            torch.cuda.synchronize()
            assert np.array_equal(d_x.cpu().numpy(), xg)
        s.close()


def test_split_projection_gauges_each_component():
    _print("case split projection")
    kw, kfun, sides, coarse = SPLIT_CASES["ring_amr"]
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    k = kfun(om)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(pm.nfaces)
    p0 = rng.standard_normal(pm.ncells)
    os_ = OSolver(om, k, coarse_size=coarse)
    uo, po, phio, so = os_.project(u, p0, 0.1, tolerance=1e-12, max_cycles=300)
    s = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), coarse_size=coarse)
    u2, p2 = u.copy(), p0.copy()
    st, phi = tpamr.project(u2, p2, s, 0.1, tolerance=1e-12, max_cycles=300)
    assert st.converged and bool(st.compatible) == bool(so["compatible"])
    assert rel(phi, phio) < 1e-10
    assert np.allclose(u2, uo, rtol=0, atol=1e-10 * np.max(np.abs(uo)))
    assert np.allclose(p2, po, rtol=0, atol=1e-10 * np.max(np.abs(po)))
    slot, roots = s.level_components(0)
    V = pm.volumes
    for q in range(1, len(roots) + 1):
        c = slot == q
        assert abs(np.dot(V[c], p2[c])) <= 1e-12 * np.sum(V[c]) * np.max(np.abs(p2))
    s.close()
