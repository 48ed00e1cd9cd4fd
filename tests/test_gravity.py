"""gravity! (src/two_phase/forces.jl:528-580): potential g . x at the cell
centres (interface centroids in mixed cells), face jumps rho_f |f| (phi+ -
phi-) with the open-face phase mixture, accelerations through
_weighted_jump_force!.

The oracle is pinned by
- the reference's hydrostatic-balance test (test/two_phase/forces.jl:528-558):
  on three meshes and two gravity directions, projecting dt * gravity! of a
  flat interface normal to g leaves max |u| <= 2048 eps (the geometry here is
  exact for a straight interface: cell fractions, phase face apertures and
  interface centroids);
- closed forms that the balance test leaves open (it holds for any sign and
  any face density): with one density on a uniform mesh the acceleration is
  g_axis on every face; on AMR meshes it is rho * weighted_gradient(g . x);
  with mixed apertures each jump is (rho_l a_l + rho_g a_g) |f| (phi+ - phi-)
  with the plus cell on the +axis side.
The GPU jumps are then bit-exact against the oracle and the accelerations
within 1e-13."""
import numpy as np
import pytest

import tpamr
from tpamr import workloads as W
from oracle import OMesh, OSolver

EPS = np.finfo(float).eps


def line_interface(gravity):
    """The reference test's level sets: y = 0.43 (liquid below) for g = (0, -1),
    and g . (x - 0.5, y - 0.5) = 0 otherwise (liquid where g . (x - c) < 0)."""
    gx, gy = gravity
    if gx == 0:
        return (lambda x: np.full_like(x, 0.43)), (lambda x: np.zeros_like(x)), True
    s = gx / gy
    return (lambda x: 0.5 - s * (x - 0.5)), (lambda x: np.full_like(x, -s)), gy > 0


def hydrostatic_meshes():
    """test/two_phase/forces.jl:529-543: one 8x8 root, walls, max_level 2;
    kind 2 refines the root and its first child; kind 3 (adapt! refine near the
    interface) refines every block that the interface band touches."""
    base = dict(block_size=(8, 8), max_level=2, periodic=(False, False))
    return {
        1: lambda g: dict(base),
        2: lambda g: dict(base, refine=lambda k, o, e: k[0] == 0 or (k[0] == 1 and k[1] == 0 and k[2] == 0)),
        3: lambda g: dict(base, refine=lambda k, o, e: _touches(line_interface(g)[0], o, e)),
    }


def _touches(eta, o, e, band=0.1):
    xs = np.linspace(o[0], o[0] + e[0], 9)
    ys = eta(xs)
    return ys.max() + band >= o[1] and ys.min() - band <= o[1] + e[1]


def geometry(pm, gravity):
    eta, deta, below = line_interface(gravity)
    return W.interface_geometry(pm, eta, deta, liquid_below=below)


@pytest.mark.parametrize("kind", [1, 2, 3])
@pytest.mark.parametrize("gravity", [(0.0, -1.0), (0.6, -0.8)])
def test_oracle_hydrostatic_balance(kind, gravity):
    kw = hydrostatic_meshes()[kind](gravity)
    pm, om = tpamr.BlockTreeMesh(**kw), OMesh(**kw)
    frac, la, ga, off = geometry(pm, gravity)
    kappa = om.mixture_coefficients(frac, 1000.0, 1.0)
    s = OSolver(om, kappa)
    u = s.weighted_jump_force(om.gravity_jumps(la, ga, (1000.0, 1.0), gravity, off)) * 0.01
    assert np.max(np.abs(u)) > 1e-3  # there is a force to balance
    u, _, _, st = s.project(u, np.zeros(om.ncells), 0.01, tolerance=512 * EPS)
    assert np.max(np.abs(u)) <= 2048 * EPS, np.max(np.abs(u))


@pytest.mark.parametrize("gravity", [(0.0, -9.81), (0.6, -0.8)])
def test_oracle_single_phase_uniform_acceleration_is_gravity(gravity):
    kw = dict(base_mesh=(32, 16), block_size=(8, 8), extent=(2.0, 1.0), periodic=(False, False))
    om = OMesh(**kw)
    rng = np.random.default_rng(11)
    la = rng.uniform(0, 1, om.nfaces)
    s = OSolver(om, np.full(om.nfaces, 1 / 3.0))
    a = s.weighted_jump_force(om.gravity_jumps(la, 1 - la, (3.0, 3.0), gravity))
    pm = tpamr.BlockTreeMesh(**kw)
    expect = np.where(pm.face_axis == 0, gravity[0], gravity[1])
    assert np.allclose(a, expect, rtol=1e-13, atol=1e-13)


def test_oracle_single_phase_amr_acceleration_is_weighted_gradient():
    kw = dict(base_mesh=(16, 16), block_size=(8, 8), max_level=2, periodic=(False, False),
              refine=lambda k, o, e: (o[0] - 0.4) ** 2 + (o[1] - 0.6) ** 2 < 0.08)
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    g = (0.3, -2.0)
    s = OSolver(om, np.full(om.nfaces, 0.5))
    a = s.weighted_jump_force(om.gravity_jumps(np.ones(om.nfaces), np.zeros(om.nfaces), (2.0, 7.0), g))
    x, y = pm.cell_centers()
    ref = 2.0 * s.weighted_gradient(g[0] * x + g[1] * y)
    assert np.allclose(a, ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref)))


def test_oracle_jumps_mixture_and_orientation():
    kw = dict(base_mesh=(16, 16), block_size=(8, 8), max_level=1, periodic=(False, False),
              refine=lambda k, o, e: k[1] == 0 and k[2] == 0)
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    rng = np.random.default_rng(12)
    la, ga = rng.uniform(0, 1, om.nfaces), rng.uniform(0, 1, om.nfaces)
    off = rng.uniform(-0.5, 0.5, (om.ncells, 2))
    g = (1.5, -0.25)
    j = om.gravity_jumps(la, ga, (1000.0, 1.2), g, off)
    x, y = pm.cell_centers()
    hx, hy = pm.cell_spacing()
    phi = g[0] * (x + off[:, 0] * hx) + g[1] * (y + off[:, 1] * hy)
    m, p = pm.face_minus, pm.face_plus
    rho = 1000.0 * la + 1.2 * ga
    assert np.allclose(j, rho * pm.face_area * (phi[p] - phi[m]), rtol=1e-13, atol=1e-13)
    # plus cell on the +axis side
    assert np.all(np.where(pm.face_axis == 0, x[p] - x[m], y[p] - y[m]) > 0)


def test_oracle_rejects_gravity_along_periodic_axis():
    om = OMesh(base_mesh=(8, 8), block_size=(8, 8), periodic=(True, False))
    with pytest.raises(Exception, match="periodic"):
        om.gravity_jumps(np.ones(om.nfaces), np.zeros(om.nfaces), (1.0, 1.0), (1.0, 0.0))
    om.gravity_jumps(np.ones(om.nfaces), np.zeros(om.nfaces), (1.0, 1.0), (0.0, -1.0))


def test_straight_interface_geometry_is_exact():
    """The interface data of a straight line (the hydrostatic inputs): cell
    fractions sum to the liquid area, apertures of faces inside one phase are
    0 or 1, centroids lie on the line."""
    pm = tpamr.BlockTreeMesh(base_mesh=(16, 16), block_size=(8, 8), periodic=(False, False), max_level=1,
                             refine=lambda k, o, e: k[1] == 0)
    eta, deta, below = line_interface((0.6, -0.8))
    frac, la, ga, off = W.interface_geometry(pm, eta, deta, liquid_below=below)
    # liquid above y = 0.5 + 0.75 (x - 0.5) in the unit square: area 1/2
    assert abs(np.sum(frac * pm.volumes) - 0.5) < 1e-14
    x, y = pm.cell_centers()
    hx, hy = pm.cell_spacing()
    mixed = (frac > 0) & (frac < 1)
    cx, cy = x + off[:, 0] * hx, y + off[:, 1] * hy
    assert np.allclose(cy[mixed], eta(cx[mixed]), atol=1e-14)
    assert np.all(la + ga == 1.0)


# GPU ------------------------------------------------------------------------

def wave_case(levels=(1, 4), block=8):
    pm = W.wave_mesh(levels, block=block)
    om = OMesh(base_mesh=pm.base_mesh, block_size=pm.block_size, extent=pm.extent, periodic=pm.periodic,
               max_level=pm.max_level, refined_keys=pm.refined_keys)
    eta, deta = W.wave_surface()
    frac, la, ga, off = W.interface_geometry(pm, eta, deta)
    return pm, om, frac, la, ga, off


@pytest.mark.gpu
def test_gravity_jumps_bit_exact_and_accelerations_match_oracle():
    pm, om, frac, la, ga, off = wave_case()
    assert np.count_nonzero(off) > 0
    kappa = om.mixture_coefficients(frac, *W.WAVE_DENSITY)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, kappa), precision="fp64", device=0)
    for o in (off, None):
        jg = tpamr.gravity_jumps(gs, la, ga, density=W.WAVE_DENSITY, gravity=W.WAVE_GRAVITY, centroid_offsets=o)
        jo = om.gravity_jumps(la, ga, W.WAVE_DENSITY, W.WAVE_GRAVITY, o)
        assert np.array_equal(jg, jo)
    ag = tpamr.gravity(gs, la, ga, density=W.WAVE_DENSITY, gravity=W.WAVE_GRAVITY, centroid_offsets=off)
    ao = OSolver(om, kappa).weighted_jump_force(om.gravity_jumps(la, ga, W.WAVE_DENSITY, W.WAVE_GRAVITY, off))
    assert np.max(np.abs(ag - ao)) <= 1e-13 * np.max(np.abs(ao))
    with pytest.raises(ValueError):
        tpamr.gravity_jumps(gs, la, ga, density=(1000.0, -1.0), gravity=W.WAVE_GRAVITY)
    gs.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("kind", [2, 3])
@pytest.mark.parametrize("gravity", [(0.0, -1.0), (0.6, -0.8)])
def test_hydrostatic_balance_on_gpu(precision, kind, gravity):
    kw = hydrostatic_meshes()[kind](gravity)
    pm = tpamr.BlockTreeMesh(**kw)
    frac, la, ga, off = geometry(pm, gravity)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, np.ones(pm.nfaces)), precision=precision, device=0)
    tpamr.refresh_two_phase(gs, frac, (1000.0, 1.0))
    u = tpamr.gravity(gs, la, ga, density=(1000.0, 1.0), gravity=gravity, centroid_offsets=off) * 0.01
    assert np.max(np.abs(u)) > 1e-3
    st, _ = tpamr.project(u, np.zeros(pm.ncells), gs, 0.01, tolerance=512 * EPS)
    assert np.max(np.abs(u)) <= 2048 * EPS, (np.max(np.abs(u)), st)
    gs.close()


@pytest.mark.gpu
def test_device_twins_of_mixture_and_gravity_match_host_entry_points():
    """tpamr_mixture_face_coefficients_device and tpamr_gravity_jumps_device
    run the host entry points' kernels on device memory: bit for bit the same
    coefficients and jumps."""
    import torch
    pm, om, frac, la, ga, off = wave_case()
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, np.ones(pm.nfaces)), precision="fp64", device=0)
    dev = torch.device("cuda:0")
    kd = tpamr.mixture_face_coefficients_device(gs, torch.as_tensor(frac, device=dev), W.WAVE_DENSITY)
    assert np.array_equal(kd.cpu().numpy(), tpamr.mixture_face_coefficients(gs, frac, W.WAVE_DENSITY))
    for o in (off, None):
        jd = tpamr.gravity_jumps_device(gs, torch.as_tensor(la, device=dev), torch.as_tensor(ga, device=dev),
                                        density=W.WAVE_DENSITY, gravity=W.WAVE_GRAVITY,
                                        centroid_offsets=None if o is None else torch.as_tensor(o, device=dev))
        jh = tpamr.gravity_jumps(gs, la, ga, density=W.WAVE_DENSITY, gravity=W.WAVE_GRAVITY, centroid_offsets=o)
        assert np.array_equal(jd.cpu().numpy(), jh)
    with pytest.raises(ValueError):
        tpamr.gravity_jumps_device(gs, torch.as_tensor(la, device=dev), torch.as_tensor(ga, device=dev),
                                   density=(1000.0, -1.0), gravity=W.WAVE_GRAVITY)
    with pytest.raises(ValueError):
        tpamr.mixture_face_coefficients_device(gs, torch.as_tensor(frac[:-1], device=dev), W.WAVE_DENSITY)
    gs.close()
