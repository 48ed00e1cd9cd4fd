"""The library's host hierarchy builder against the oracle (CPU, host-only
context): identical block and face schedules, multigrid level sizes, Jacobi
weights, level operators, tentative/smoothed transfers and restricted face
coefficients."""
import os
import subprocess
import sys

import numpy as np
import pytest

import tpamr
from cases import MESHES, near_circle
from oracle import OMesh, OSolver

CASES = [("uniform64", "one"), ("periodic_tg", "one"), ("amr32", "one"), ("amr64", "one"),
         ("amr_mixed", "smooth"), ("pruned_tree", "smooth"), ("circle16", "smooth"), ("amr_block4", "smooth"),
         ("amr_block32", "smooth")]


def kappa(m, kind):
    return np.ones(m.nfaces) if kind == "one" else np.exp(0.4 * np.sin(np.arange(1, m.nfaces + 1)))


@pytest.mark.parametrize("name,kind", CASES)
def test_hierarchy_matches_oracle(name, kind):
    kw = MESHES[name]
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    assert (om.ncells, om.nfaces, om.nbfaces) == (pm.ncells, pm.nfaces, pm.nboundary_faces)
    assert np.array_equal(om.block_keys, pm.block_keys)
    for a, b in [(om.face_minus, pm.face_minus), (om.face_plus, pm.face_plus), (om.face_axis, pm.face_axis),
                 (om.face_area, pm.face_area), (om.bface_cell, pm.bface_cell), (om.bface_side, pm.bface_side)]:
        assert np.array_equal(a, b)
    k = kappa(pm, kind)
    os_, ph = OSolver(om, k), tpamr.HostHierarchy(pm, k)
    assert [os_.level_info(l)["ncells"] for l in range(os_.nlevels)] == ph.sizes()
    for l in range(os_.nlevels):
        A, B = os_.level_matrix(l).tocsr(), ph.level_matrix(l)
        assert abs(A - B).max() <= 1e-15 * abs(A).max()
        assert abs(os_.level_info(l)["omega"] - ph.level_info(l)["omega"]) <= 1e-15
        assert np.allclose(os_.level_kappa(l), ph.level_kappa(l), rtol=1e-14, atol=0)
        if l + 1 < os_.nlevels:
            assert abs(os_.level_matrix(l, "P").tocsr() - ph.prolongation(l)).max() <= 1e-15


def test_invalid_inputs_raise():
    with pytest.raises(ValueError):
        tpamr.BlockTreeMesh(block_size=(6, 6), base_mesh=(12, 9))
    with pytest.raises(ValueError):
        tpamr.BlockTreeMesh(block_size=(5, 5))
    m = tpamr.BlockTreeMesh(block_size=(8, 8))
    with pytest.raises(ValueError):
        tpamr.HostHierarchy(m, -np.ones(m.nfaces))  # negative coefficients are rejected
    with pytest.raises(ValueError):
        tpamr.MultigridOptions(coarse_size=0)
    with pytest.raises(ValueError):
        tpamr.LinearSolverOptions(krylov="cg")


def test_threaded_builder_does_not_depend_on_thread_count(tmp_path):
    """The setup passes split their work by the host thread count; the faces,
    level operators, weights, restricted coefficients and transfers are the
    same, bit for bit, on 1 and on 16 threads (a mesh large enough that every
    pass runs threaded)."""
    here = os.path.dirname(__file__)
    outs = []
    for threads in (1, 16):
        path = tmp_path / f"h{threads}.npz"
        env = dict(os.environ, TPAMR_HOST_THREADS=str(threads))
        subprocess.run([sys.executable, os.path.join(here, "_host_dump.py"), str(path)], env=env, check=True,
                       cwd=here, timeout=600)
        outs.append(np.load(path))
    a, b = outs
    assert sorted(a.files) == sorted(b.files)
    for k in a.files:
        assert np.array_equal(a[k], b[k]), k


def _circle_mesh(centre):
    return dict(base_mesh=(16, 16), block_size=(16, 16), periodic=(False, False), max_level=4,
                refine=lambda k, o, e: k[0] < 2 or near_circle(o, e, centre=centre, radius=0.15, band=4.0 / 256))


@pytest.mark.parametrize("seq", [[(0.5, 0.6), (0.53, 0.58)], [(0.5, 0.6), (0.3, 0.3), (0.5, 0.6)]])
def test_hierarchy_reuse_after_remesh(seq):
    """A context that changes mesh takes over the coarse levels whose tree
    topology is unchanged (_tree_reuse_suffix!, tree_coarsening.jl:214-226):
    the hierarchy equals a fresh build of the new mesh and the oracle's."""
    h = None
    for centre in seq:
        kw = _circle_mesh(centre)
        pm = tpamr.BlockTreeMesh(**kw)
        k = kappa(pm, "smooth")
        if h is None:
            h = tpamr.HostHierarchy(pm, k)
            assert h.reuse() == (-1, -1)
            continue
        h.remesh(pm, k)
        new_start, old_start = h.reuse()
        assert new_start > 0 and old_start >= 0
        fresh = tpamr.HostHierarchy(pm, k)
        os_ = OSolver(OMesh(**kw), k)
        assert h.sizes() == fresh.sizes() == [os_.level_info(l)["ncells"] for l in range(os_.nlevels)]
        for l in range(h.nlevels):
            A, B = h.level_matrix(l), fresh.level_matrix(l)
            assert (A != B).nnz == 0
            assert abs(os_.level_matrix(l).tocsr() - A).max() <= 1e-15 * abs(A).max()
            assert np.array_equal(h.level_kappa(l), fresh.level_kappa(l))
            assert h.level_info(l) == fresh.level_info(l)
            if l + 1 < h.nlevels:
                assert np.array_equal(h.level_transfer(l), fresh.level_transfer(l))
