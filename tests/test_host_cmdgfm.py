"""The library's CMDGFM block-system setup on a host-only context against the
oracle (oracle/oracle_cmdgfm.py, pinned in tests/test_oracle_cmdgfm.py): the
jump Schur complement and its Crout ILU are bit-identical for several drop
tolerances, and malformed systems raise."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_cmdgfm as OC
import tpamr
from cmdgfm_cases import build_system
from oracle import OMesh, OSolver

KW = dict(base_mesh=(32, 32), block_size=(8, 8), periodic=(False, False))


def _case(kw=KW, seed=0):
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    cx, cy = om.cell_centers()
    rho = np.where(np.hypot(cx - 0.5, cy - 0.5) < 0.27, 1000.0, 1.0)
    kap = 2.0 / (rho[om.face_minus] + rho[om.face_plus])
    osol = OSolver(om, kap, coarse_size=16)
    d = build_system(om, osol, seed=seed)
    return d, tpamr.HostHierarchy(pm, kap, tpamr.MultigridOptions(coarse_size=16))


def _system(h, d, **kw):
    return tpamr.CMDGFMBlockSystem(h, d["matrix"], d["pressure_indices"], d["pressure_anchors"],
                                   d["pressure_volumes"], d["jump_lengths"], omitted=d["omitted"],
                                   omitted_rhs=d["omitted_rhs"], omitted_volumes=d["omitted_volumes"], **kw)


def _same(a, b):
    a, b = sp.csr_matrix(a), sp.csr_matrix(b)
    a.sort_indices()
    b.sort_indices()
    return (a.shape == b.shape and np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
            and np.array_equal(a.data, b.data))


@pytest.mark.parametrize("tau", [0.0, 1e-3, 0.05])
def test_jump_factor_equals_oracle(tau):
    d, h = _case()
    sysm = _system(h, d, drop_tolerance=tau)
    s, lo, up = sysm.jump_factor()
    s_ref = OC.jump_schur(d["matrix"], d["np"])
    lo_ref, up_ref = OC.crout_ilu(s_ref, tau)
    assert _same(s, s_ref)
    assert _same(lo, lo_ref)
    assert _same(up, up_ref)
    if tau > 0:
        assert lo.nnz + up.nnz < OC.crout_ilu(s_ref, 0.0)[0].nnz + OC.crout_ilu(s_ref, 0.0)[1].nnz


def test_system_without_jumps_has_no_factor():
    om, pm = OMesh(**KW), tpamr.BlockTreeMesh(**KW)
    osol = OSolver(om, 1.0, coarse_size=16)
    d = build_system(om, osol, jumps=False)
    h = tpamr.HostHierarchy(pm, 1.0)
    s, lo, up = _system(h, d).jump_factor()
    assert s.shape == (0, 0)


def test_malformed_systems_raise():
    d, h = _case()
    idx = d["pressure_indices"].copy()
    dup = idx.copy()
    dup[1] = dup[0]
    with pytest.raises(ValueError, match="distinct"):
        tpamr.CMDGFMBlockSystem(h, d["matrix"], dup, d["pressure_anchors"], d["pressure_volumes"],
                                d["jump_lengths"])
    bad_anchor = d["pressure_anchors"].copy()
    bad_anchor[0] = idx[5]
    with pytest.raises(ValueError, match="anchor"):
        tpamr.CMDGFMBlockSystem(h, d["matrix"], idx, bad_anchor, d["pressure_volumes"], d["jump_lengths"])
    with pytest.raises(ValueError, match="outside"):
        tpamr.CMDGFMBlockSystem(h, d["matrix"], np.where(idx == idx[3], 10 ** 6, idx), d["pressure_anchors"],
                                d["pressure_volumes"], d["jump_lengths"])
    zero_diag = d["matrix"].tolil()
    zero_diag[0, 0] = 0.0
    with pytest.raises(ValueError, match="diagonal"):
        tpamr.CMDGFMBlockSystem(h, sp.csr_matrix(zero_diag), idx, d["pressure_anchors"], d["pressure_volumes"],
                                d["jump_lengths"])
    with pytest.raises(ValueError, match="jump lengths"):
        tpamr.CMDGFMBlockSystem(h, d["matrix"], idx, d["pressure_anchors"], d["pressure_volumes"],
                                np.zeros(d["nx"]))
    with pytest.raises(ValueError):
        _system(h, d, drop_tolerance=-1.0)
    sysm = _system(h, d)
    with pytest.raises(RuntimeError, match="host-only"):
        sysm.solve(d["rhs"], dt=d["dt"], density=d["density"])
