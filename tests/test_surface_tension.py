"""Surface-tension jumps with the face curvature of `_face_curvature`
(src/two_phase/forces.jl:211-300, 345-378).

The oracle is pinned by the reference's own known answers: the synthetic
two-cell data of test/two_phase/forces.jl:487-520 (no height-function cell:
the c(1-c) interpolant bit for bit; one: its curvature; two: the interpolant;
continuity in the coverage and the fractions; symmetry; a convex
combination) and the exhaustive-search check of the curvature extension
(test/two_phase/forces.jl:390-433: on a two-level mesh, walls and periodic,
the block-culled search equals a search over every cell, bit for bit). The
GPU jumps are then compared with the oracle's: bit for bit where a face has a
mixed neighbour, and to 1e-14 through the extension search (whose distance
is `hypot`, rounded by the device library)."""
import numpy as np
import pytest

import tpamr
from cases import disk_fraction
from oracle import OMesh, OSolver

EPS = np.finfo(float).eps


def mixed_face_mesh():
    kw = dict(base_mesh=(64, 64), block_size=(8, 8), periodic=(True, True))
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    return om, pm


def curvature_of(om, pm, face):
    m, p = pm.face_minus[face], pm.face_plus[face]
    c = np.zeros(pm.ncells)
    k = np.full(pm.ncells, np.nan)
    q = np.zeros(pm.ncells)

    def f(cm, km, qm, cp, kp, qp):
        c[m], k[m], q[m], c[p], k[p], q[p] = cm, km, qm, cp, kp, qp
        return om.face_curvature(c, k, q, [face])[0]

    return f


def interpolant(cm, km, cp, kp):
    return (cm * (1 - cm) * km + cp * (1 - cp) * kp) / (cm * (1 - cm) + cp * (1 - cp))


def test_oracle_lhf_first_face_curvature_known_answers():
    om, pm = mixed_face_mesh()
    curvature = curvature_of(om, pm, 100)
    assert curvature(0.3, 5.0, 0.0, 0.01, 7.0, 0.0) == interpolant(0.3, 5.0, 0.01, 7.0)
    assert abs(curvature(0.3, 5.0, 1.0, 0.01, 7.0, 0.0) - 5.0) <= 2 * EPS * 5.0
    assert abs(curvature(0.01, 5.0, 1.0, 0.3, 7.0, 0.0) - 5.0) <= 2 * EPS * 5.0
    assert abs(curvature(0.01, 7.0, 0.0, 0.3, 5.0, 1.0) - 5.0) <= 2 * EPS * 5.0
    ref = interpolant(0.3, 5.0, 0.6, 7.0)
    assert abs(curvature(0.3, 5.0, 1.0, 0.6, 7.0, 1.0) - ref) <= 4 * EPS * ref
    assert abs(curvature(0.3, 5.0, 1.0e-9, 0.01, 7.0, 0.0) - interpolant(0.3, 5.0, 0.01, 7.0)) <= 1.0e-8
    assert abs(curvature(0.3, 5.0, 1 - 1.0e-9, 0.01, 7.0, 0.0) - 5.0) <= 1.0e-8
    assert abs(curvature(0.01, 5.0, 1 - 1.0e-9, 0.3, 7.0, 0.0) - 5.0) <= 1.0e-7
    assert abs(curvature(0.3 + 1.0e-9, 5.0, 0.6, 0.2, 7.0, 0.1) - curvature(0.3, 5.0, 0.6, 0.2, 7.0, 0.1)) <= 1.0e-8
    assert abs(curvature(1.0e-10, 5.0, 0.4, 0.3, 7.0, 0.0) - 7.0) <= 1.0e-8
    for cm, km, qm, cp, kp, qp in ((0.3, 5.0, 0.7, 0.2, 7.0, 0.1), (0.01, -2.0, 0.0, 0.45, 3.0, 0.96),
                                   (0.9, 1.0, 0.5, 0.05, 4.0, 1.0), (0.6, 2.0, 0.2, 0.6, 3.0, 0.2)):
        value = curvature(cm, km, qm, cp, kp, qp)
        assert value == curvature(cp, kp, qp, cm, km, qm)
        assert min(km, kp) <= value <= max(km, kp)


def wrapped(d, extent, periodic):
    return d - np.rint(d / extent) * extent if periodic else d


@pytest.mark.parametrize("periodic", [(False, False), (True, True)])
def test_oracle_curvature_extension_equals_exhaustive_search(periodic):
    kw = dict(base_mesh=(16, 16), block_size=(4, 4), max_level=1, periodic=periodic,
              refine=lambda k, o, e: k[0] == 0 and k[1] == 0)
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    x, y = om.cell_centers()
    hx, hy = pm.cell_spacing()
    c = disk_fraction(x, y, hx, hy, (0.15, 0.5), 0.24)
    k = 1 / 0.24 + 0.3 * np.sin(np.arange(pm.ncells))
    q = np.zeros(pm.ncells)
    h = np.maximum(hx, hy)
    sample = (c > 0) & (c < 1) & np.isfinite(k)
    fx, fy = om.face_center[:, 0], om.face_center[:, 1]
    pure = np.nonzero(np.isin(c[pm.face_minus], (0.0, 1.0)) & np.isin(c[pm.face_plus], (0.0, 1.0)))[0]
    got = om.face_curvature(c, k, q, pure)
    checked = 0
    for n, f in enumerate(pure):
        m, p = pm.face_minus[f], pm.face_plus[f]
        radius = 3 * max(h[m], h[p])
        total = weight = 0.0
        for cell in np.nonzero(sample)[0]:
            dx = wrapped(x[cell] - fx[f], 1.0, periodic[0])
            dy = wrapped(y[cell] - fy[f], 1.0, periodic[1])
            d = np.hypot(dx, dy)
            if not d <= radius + h[cell]:
                continue
            gw = c[cell] * (1 - c[cell]) / max(d * d, (h[cell] / 4) ** 2)
            total += gw * k[cell]
            weight += gw
        if weight > 0:
            assert got[n] == total / weight, f
            checked += 1
        else:
            assert np.isnan(got[n])
    assert checked > 0


def test_oracle_cell_jumps_use_face_curvature():
    om, pm = mixed_face_mesh()
    x, y = pm.cell_centers()
    hx, hy = pm.cell_spacing()
    c = disk_fraction(x, y, hx, hy, (0.5, 0.5), 0.2)
    k = np.where((c > 0) & (c < 1), 5.0 + 0.01 * np.cos(np.arange(pm.ncells)), np.nan)
    q = np.where(np.arange(pm.ncells) % 3 == 0, 1.0, 0.25)
    for csf in (False, True):
        j, unresolved = om.surface_tension_jumps_cells(c, k, q, 0.072, csf=csf)
        kf = om.face_curvature(c, k, q, np.arange(pm.nfaces))
        jump = (c[pm.face_plus] - c[pm.face_minus]) if csf else \
            (c[pm.face_plus] > 0.5).astype(float) - (c[pm.face_minus] > 0.5)
        on = jump != 0
        assert np.array_equal(j[on & ~np.isnan(kf)], (0.072 * kf * jump * pm.face_area)[on & ~np.isnan(kf)])
        assert np.all(j[~on] == 0)
        assert unresolved == np.count_nonzero(on & np.isnan(kf))


@pytest.mark.gpu
@pytest.mark.parametrize("csf", [False, True])
def test_cell_curvature_jumps_on_gpu_match_oracle(csf):
    kw = dict(base_mesh=(32, 32), block_size=(8, 8), max_level=1, periodic=(False, False),
              refine=lambda k, o, e: k[1] < 2)
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    x, y = pm.cell_centers()
    hx, hy = pm.cell_spacing()
    c = disk_fraction(x, y, hx, hy, (0.6, 0.45), 0.21)
    # a flat interface on the grid line y = 0.75 (pure cells on both sides:
    # the extension search) next to a few mixed cells with curvatures
    c = np.where(y > 0.75, 0.0, np.where(y > 0.7, 1.0, c))
    band = (y > 0.74) & (y < 0.8) & (np.arange(pm.ncells) % 5 == 0)
    c = np.where(band & (y > 0.75), 0.3, c)
    k = np.where((c > 0) & (c < 1), 1 / 0.21 + 0.05 * np.sin(np.arange(pm.ncells)), np.nan)
    q = np.where(np.arange(pm.ncells) % 4 == 0, 0.0, 0.8)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, np.ones(pm.nfaces)), precision="fp64", device=0)
    jg, ng = tpamr.surface_tension_jumps_from_cells(gs, c, k, q, sigma=0.072,
                                                    formulation=tpamr.CSF if csf else tpamr.CGFM)
    jo, no = om.surface_tension_jumps_cells(c, k, q, 0.072, csf=csf)
    gs.close()
    assert ng == no
    m, p = pm.face_minus, pm.face_plus
    cm, cp = c[m], c[p]
    direct = ((cm > 0) & (cm < 1)) | ((cp > 0) & (cp < 1))
    assert np.array_equal(jg[direct], jo[direct])
    assert np.max(np.abs(jg - jo)) <= 1e-14 * np.max(np.abs(jo))
    extended = ~direct & (jo != 0)
    print("faces through the extension search:", int(np.count_nonzero(extended)), "unresolved:", ng, flush=True)
    assert np.count_nonzero(extended) > 0
