"""Halo refresh after the projection: fill_halo! (src/mesh/variables.jl:205-229)
with the schedule of _build_halo_schedule (src/mesh/topology/mesh.jl:944-1037).

The oracle is pinned by the reference's own known answers
(test/mesh/variables.jl:5-53: a constant field fills every halo with the
constant on three meshes; on the two-level mesh the coarse block's halo column
beside finer leaves equals the linear field at the coarse halo centre, the
conservative average; Dirichlet values land on every physical halo) and by
invariants that a wrong index, weight or wrap breaks (a same-level halo is
the neighbour's interior value bit for bit; periodic wrap; halo width 2; the
schedule does not depend on origin and extent, test/mesh/topology.jl:393-466).
The library's schedule (host-built) is then compared with the oracle on the
CPU, and the GPU gather with the oracle bit for bit."""
import numpy as np
import pytest

import tpamr
from cases import MESHES
from oracle import OMesh


def ref_meshes():
    """The meshes of test/mesh/variables.jl:6-17 (one 4x4 root, periodic)."""
    return [dict(block_size=(4, 4)),
            dict(block_size=(4, 4), max_level=1, refine=lambda k, o, e: True),
            dict(block_size=(4, 4), max_level=2,
                 refine=lambda k, o, e: k[0] == 0 or (k[0] == 1 and k[1] == 0 and k[2] == 0))]


def block_index(m, key):
    return [q for q in range(m.nblocks) if tuple(m.block_keys[q]) == key][0]


@pytest.mark.parametrize("kw", ref_meshes())
def test_oracle_constant_field(kw):
    m = OMesh(**kw)
    out = m.fill_halo(np.full(m.ncells, 3.25))
    assert np.all(out == 3.25)


def test_oracle_coarse_halo_beside_finer_leaves_is_conservative_average():
    m = OMesh(**ref_meshes()[2])
    x, _ = m.cell_centers()
    out = m.fill_halo(x)
    b = block_index(m, (1, 1, 0))
    ox, spacing = m.block_geom[b][0], m.block_geom[b][4]
    # west halo column (i = 0 in the haloed array), rows 1..4
    assert np.allclose(out[b, 1:5, 0], ox - spacing / 2, rtol=0, atol=1e-15)


def test_oracle_dirichlet_physical_halos():
    m = OMesh(block_size=(4, 4), periodic=(False, False))
    x, y = m.cell_centers()
    m.fill_halo(x + 2 * y)
    out = m.fill_halo(x + 2 * y, physical=np.full(m.physical_halo_entries, 7.0))
    assert m.physical_halo_entries == 6 * 6 - 16
    blk = out[0]
    assert np.all(blk[0, :] == 7.0) and np.all(blk[-1, :] == 7.0)
    assert np.all(blk[:, 0] == 7.0) and np.all(blk[:, -1] == 7.0)


def test_oracle_same_level_and_periodic_halos_copy_neighbours():
    m = OMesh(base_mesh=(16, 16), block_size=(8, 8), periodic=(True, True))
    rng = np.random.default_rng(3)
    v = rng.standard_normal(m.ncells)
    out = m.fill_halo(v)
    grid = v.reshape(2, 2, 8, 8).transpose(0, 2, 1, 3).reshape(16, 16)  # blocks (j, i) in (level, j, i) order
    for b in range(m.nblocks):
        bj, bi = m.block_keys[b][2], m.block_keys[b][1]
        full = np.roll(np.roll(grid, 1 - 8 * bj, axis=0), 1 - 8 * bi, axis=1)[:10, :10]
        assert np.array_equal(out[b], full)


def test_oracle_zero_gradient_walls_copy_nearest_interior():
    m = OMesh(base_mesh=(8, 8), block_size=(8, 8), periodic=(False, False))
    v = np.random.default_rng(4).standard_normal(m.ncells).reshape(8, 8)
    out = m.fill_halo(v.ravel())[0]
    assert np.array_equal(out[1:9, 1:9], v)
    assert np.array_equal(out[0, 1:9], v[0]) and np.array_equal(out[9, 1:9], v[7])
    assert np.array_equal(out[1:9, 0], v[:, 0]) and np.array_equal(out[1:9, 9], v[:, 7])
    assert out[0, 0] == v[0, 0] and out[9, 9] == v[7, 7]


@pytest.mark.parametrize("halo", [1, 2])
def test_oracle_schedule_is_independent_of_geometry(halo):
    kw = dict(base_mesh=(16, 8), block_size=(4, 4), max_level=2, periodic=(True, False),
              refine=lambda k, o, e: k[0] == 0 and k[1] < 2)
    a = OMesh(**kw)
    b = OMesh(origin=(-3.0, 5.0), extent=(7.0, 0.25), **kw)
    v = np.random.default_rng(5).standard_normal(a.ncells)
    assert np.array_equal(a.fill_halo(v, halo), b.fill_halo(v, halo))


def apply_schedule(S, cells, nblocks, bx, by, halo, physical=None):
    """Gather of the library's schedule on the CPU, donors summed in order."""
    W, H = bx + 2 * halo, by + 2 * halo
    out = np.zeros(nblocks * W * H)
    c = np.arange(nblocks * bx * by)
    blk, loc = c // (bx * by), c % (bx * by)
    out[blk * W * H + (loc // bx + halo) * W + (loc % bx + halo)] = cells
    ptr, donor, weight = S["ptr"], S["donor"], S["weight"]
    cnt = np.diff(ptr)
    v = np.zeros(len(cnt))
    for k in range(4):
        m = cnt > k
        v[m] += weight[ptr[:-1][m] + k] * cells[donor[ptr[:-1][m] + k]]
    ph = S["phys"] >= 0
    v[ph] = cells[donor[ptr[:-1][ph]]] if physical is None else physical[S["phys"][ph]]
    out[S["dest"]] = v
    return out.reshape(nblocks, H, W)


SCHEDULE_CASES = [("amr_mixed", 1), ("circle16", 1), ("pruned_tree", 2), ("amr32", 2), ("periodic_tg", 1)]


def meshes_for(name):
    kw = MESHES[name]
    return tpamr.BlockTreeMesh(**kw), OMesh(**kw)


@pytest.mark.parametrize("name,halo", SCHEDULE_CASES)
def test_library_schedule_matches_oracle(name, halo):
    pm, om = meshes_for(name)
    h = tpamr.HostHierarchy(pm, 1.0)
    S = tpamr.halo_schedule(h.ctx, halo)
    v = np.random.default_rng(6).standard_normal(pm.ncells)
    bx, by = pm.block_size
    assert np.array_equal(apply_schedule(S, v, pm.nblocks, bx, by, halo), om.fill_halo(v, halo))
    assert S["nphysical"] == om.physical_halo_entries
    if S["nphysical"]:
        phys = np.random.default_rng(7).standard_normal(S["nphysical"])
        assert np.array_equal(apply_schedule(S, v, pm.nblocks, bx, by, halo, phys), om.fill_halo(v, halo, phys))
    # weights of every entry sum to one; donors are interior cells
    sums = np.add.reduceat(S["weight"], S["ptr"][:-1])
    assert np.allclose(sums, 1.0, rtol=0, atol=0) and S["donor"].min() >= 0 and S["donor"].max() < pm.ncells


@pytest.mark.gpu
@pytest.mark.parametrize("name,halo", SCHEDULE_CASES)
def test_fill_halo_on_gpu_matches_oracle(name, halo):
    print("case fill_halo", name, halo, flush=True)
    pm, om = meshes_for(name)
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, 1.0))
    v = np.random.default_rng(8).standard_normal(pm.ncells)
    assert np.array_equal(tpamr.fill_halo(gs, v, halo), om.fill_halo(v, halo))
    n = om.physical_halo_entries
    if n:
        phys = np.linspace(-1.0, 2.0, n)
        assert np.array_equal(tpamr.fill_halo(gs, v, halo, physical=phys), om.fill_halo(v, halo, phys))
    gs.close()
