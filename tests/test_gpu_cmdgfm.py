"""GPU: the CMDGFM reduced-pressure/jump block solve through the C ABI against
the CPU FP64 oracle (oracle/oracle_cmdgfm.py, pinned in
tests/test_oracle_cmdgfm.py) on assembled systems of the reference's
structure (tests/cmdgfm_cases.py): walled, periodic and AMR pressure meshes,
three pressure components, no jumps, warm starts, the unpreconditioned
solve, the device entry point and the FP32 V-cycle.

Bar: the same BiCGSTAB iteration and Krylov-run counts and warm-start
decision as the oracle, every defect within the tolerance, and the solution
within 1e-6 (relative) of the oracle's; the device entry is bit-identical to
the host-array entry.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle_cmdgfm as OC  # noqa: E402
import tpamr  # noqa: E402
from cases import MESHES  # noqa: E402
from cmdgfm_cases import build_system  # noqa: E402
from oracle import OMesh, OSolver  # noqa: E402

TOL = 1e-8


def _print(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _kappa(om, strips=False):
    if strips:
        k = np.ones(om.nfaces)
        k[(om.face_axis == 0) & np.isin(om.face_center[:, 0], [0.25, 0.75])] = 0.0
        return k
    cx, cy = om.cell_centers()
    rho = np.where(np.hypot(cx - 0.5, cy - 0.5) < 0.27, 1000.0, 1.0)
    return 2.0 / (rho[om.face_minus] + rho[om.face_plus])


CASES = {
    "uniform32": (dict(base_mesh=(32, 32), block_size=(8, 8), periodic=(False, False)), False, True),
    "periodic64": (dict(base_mesh=(64, 64), block_size=(16, 16), periodic=(True, True)), False, True),
    "amr32": (MESHES["amr32"], False, True),
    "strips16": (dict(base_mesh=(16, 16), block_size=(4, 4), periodic=(False, False)), True, True),
    "nojump32": (dict(base_mesh=(32, 32), block_size=(8, 8), periodic=(False, False)), False, False),
    "uniform256": (dict(base_mesh=(256, 256), block_size=(16, 16), periodic=(False, False)), False, True),
}


def _setup(name, precision="fp64", seed=0):
    kw, strips, jumps = CASES[name]
    om, pm = OMesh(**kw), tpamr.BlockTreeMesh(**kw)
    k = _kappa(om, strips)
    osol = OSolver(om, k, coarse_size=16)
    d = build_system(om, osol, jumps=jumps, seed=seed)
    solver = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, k), precision=precision, coarse_size=16)
    gsys = tpamr.CMDGFMBlockSystem(solver, d["matrix"], d["pressure_indices"], d["pressure_anchors"],
                                   d["pressure_volumes"], d["jump_lengths"], omitted=d["omitted"],
                                   omitted_rhs=d["omitted_rhs"], omitted_volumes=d["omitted_volumes"])
    osys = OC.BlockSystem(d["matrix"], d["np"], d["pressure_indices"], d["pressure_anchors"], d["pressure_volumes"],
                          d["jump_lengths"], d["omitted"], d["omitted_rhs"], d["omitted_volumes"], d["ncells"],
                          osol.vcycle)
    return d, solver, gsys, osys


def _oracle(osys, d, **kw):
    return OC.solve(osys, d["rhs"], dt=d["dt"], density=d["density"], reference_length=1.0, tolerance=TOL, **kw)


def _check(st, ost, x, ox):
    assert st.converged and ost["converged"]
    assert st.iterations == ost["iterations"]
    assert st.krylov_runs == ost["krylov_runs"]
    assert st.initial_guess_used == ost["initial_guess_used"]
    assert st.pressure_defect <= TOL and st.jump_defect <= TOL and st.anchor_defect <= TOL
    assert st.pressure_unknowns == ost["pressure_unknowns"] and st.jump_unknowns == ost["jump_unknowns"]
    assert st.asymmetry == ost["asymmetry"]
    assert st.jump_diagonal_margin == ost["jump_diagonal_margin"]
    assert abs(st.initial_residual - ost["initial_residual"]) <= 1e-13 * ost["initial_residual"]
    assert rel(x, ox) <= 1e-6


@pytest.mark.parametrize("name", sorted(CASES))
def test_block_solve_matches_oracle(name):
    _print("case", name)
    d, solver, gsys, osys = _setup(name)
    x, st = gsys.solve(d["rhs"], dt=d["dt"], density=d["density"], tolerance=TOL)
    ox, ost = _oracle(osys, d)
    _print(name, "np", d["np"], "nx", d["nx"], "iterations", st.iterations, ost["iterations"], "rel", rel(x, ox))
    _check(st, ost, x, ox)
    solver.close()


def test_warm_start_from_previous_increment():
    _print("case warm start uniform32")
    d, solver, gsys, osys = _setup("uniform32")
    x0, _ = gsys.solve(d["rhs"], dt=d["dt"], density=d["density"], tolerance=TOL)
    prev = np.zeros(d["ncells"])
    prev[d["pressure_indices"]] = x0[:d["np"]]
    # the next step's right-hand side: perturbed jump rows (the pressure rows
    # stay consistent with the anchors' omitted equations)
    npp = d["np"]
    d["rhs"] = d["rhs"].copy()
    d["rhs"][npp:] *= 1.0 + 0.01 * np.cos(np.arange(d["nx"]))
    x, st = gsys.solve(d["rhs"], dt=d["dt"], density=d["density"], tolerance=TOL, previous_increment=prev)
    ox, ost = _oracle(osys, d, previous_increment=prev)
    assert st.initial_guess_used and ost["initial_guess_used"]
    _check(st, ost, x, ox)
    # warm_start=False keeps the zero start
    x2, st2 = gsys.solve(d["rhs"], dt=d["dt"], density=d["density"], tolerance=TOL, warm_start=False)
    ox2, ost2 = _oracle(osys, d, warm_start=False)
    assert not st2.initial_guess_used
    _check(st2, ost2, x2, ox2)
    solver.close()


def test_unpreconditioned_solve_converges():
    _print("case unpreconditioned strips16")
    d, solver, gsys, osys = _setup("strips16")
    x, st = gsys.solve(d["rhs"], dt=d["dt"], density=d["density"], tolerance=TOL, preconditioner=None,
                       max_iterations=3000)
    ox, ost = _oracle(osys, d, preconditioned=False, max_iterations=3000)
    assert st.preconditioner == "none"
    assert st.converged and ost["converged"]
    assert st.pressure_defect <= TOL and st.jump_defect <= TOL
    assert rel(x, ox) <= 1e-5
    solver.close()


def test_device_entry_equals_host_entry():
    import torch

    _print("case device entry amr32")
    d, solver, gsys, osys = _setup("amr32")
    x, st = gsys.solve(d["rhs"], dt=d["dt"], density=d["density"], tolerance=TOL)
    rhs = torch.tensor(d["rhs"], dtype=torch.float64, device="cuda:0")
    out = torch.full((d["np"] + d["nx"],), 7.0, dtype=torch.float64, device="cuda:0")
    st_d = gsys.solve_device(rhs, out, dt=d["dt"], density=d["density"], tolerance=TOL)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), x)
    assert st_d.iterations == st.iterations and st_d.final_residual == st.final_residual
    solver.close()


def test_fp32_vcycle_pressure_block():
    _print("case fp32 V-cycle uniform32")
    d, solver, gsys, osys = _setup("uniform32", precision="fp32_vcycle")
    x, st = gsys.solve(d["rhs"], dt=d["dt"], density=d["density"], tolerance=TOL)
    ox, ost = _oracle(osys, d)
    assert st.converged
    assert st.pressure_defect <= TOL and st.jump_defect <= TOL and st.anchor_defect <= TOL
    assert abs(st.iterations - ost["iterations"]) <= 1
    assert rel(x, ox) <= 1e-6
    solver.close()


def test_new_mesh_drops_the_system():
    _print("case set_mesh drops the system")
    d, solver, gsys, osys = _setup("nojump32")
    kw = dict(base_mesh=(16, 16), block_size=(8, 8), periodic=(False, False))
    tpamr.BlockTreeMesh(**kw).attach(solver.ctx)
    with pytest.raises(RuntimeError, match="block system"):
        gsys.solve(d["rhs"], dt=d["dt"], density=d["density"])
    solver.close()


def test_block_solve_at_size():
    """The 1024^2-cell system of tpamr.workloads.cmdgfm_block_workload
    (beyond the oracle's reach): the solve converges and the residual
    recomputed on the host from the assembled matrix meets the pressure and
    jump defects the library reports."""
    from tpamr.workloads import cmdgfm_block_workload

    _print("case cmdgfm 1024^2")
    mesh, kappa, sysd, rhs, dt, rho = cmdgfm_block_workload(n=1024)
    solver = tpamr.EllipticSolver(tpamr.PoissonOperator(mesh, kappa))
    gsys = tpamr.CMDGFMBlockSystem(solver, sysd["matrix"], sysd["pressure_indices"], sysd["pressure_anchors"],
                                   sysd["pressure_volumes"], sysd["jump_lengths"], omitted=sysd["omitted"],
                                   omitted_rhs=sysd["omitted_rhs"], omitted_volumes=sysd["omitted_volumes"])
    x, st = gsys.solve(rhs, dt=dt, density=rho, tolerance=TOL)
    solver.close()
    assert st.converged and st.jump_unknowns > 1000
    r = rhs - sysd["matrix"] @ x
    npp = len(sysd["pressure_indices"])
    pd = dt * dt * np.max(np.abs(r[:npp]) / sysd["pressure_volumes"])
    jd = dt * dt * np.max(np.abs(r[npp:]) / sysd["jump_lengths"])
    assert pd <= TOL and jd <= TOL
    assert abs(pd - st.pressure_defect) <= 1e-3 * TOL and abs(jd - st.jump_defect) <= 1e-3 * TOL
    ra = sysd["omitted_rhs"] - sysd["omitted"] @ x
    assert dt * dt * np.max(np.abs(ra) / sysd["omitted_volumes"]) <= TOL


def test_iteration_budget_and_pure_jump_system():
    """An exhausted budget stops where the oracle stops (two BiCGSTAB
    iterations, not converged, the same iterate), and a system of jumps only
    (no reduced pressures) runs the jump factor alone."""
    import scipy.sparse as sp

    _print("case budget uniform32")
    d, solver, gsys, osys = _setup("uniform32")
    x, st = gsys.solve(d["rhs"], dt=d["dt"], density=d["density"], tolerance=TOL, max_iterations=2,
                       warm_start=False)
    ox, ost = _oracle(osys, d, max_iterations=2, warm_start=False)
    assert st.iterations == ost["iterations"] == 2 and st.krylov_runs == ost["krylov_runs"] == 1
    assert not st.converged and not ost["converged"]
    assert rel(x, ox) <= 1e-10
    _print("case pure jump system")
    npp = d["np"]
    j = sp.csr_matrix(d["matrix"])[npp:, npp:]
    rj = d["rhs"][npp:]
    empty = np.zeros(0, np.int64)
    gj = tpamr.CMDGFMBlockSystem(solver, j, empty, empty, np.zeros(0), d["jump_lengths"])
    xj, sj = gj.solve(rj, dt=d["dt"], density=d["density"], tolerance=TOL)
    oj = OC.BlockSystem(j, 0, empty, empty, np.zeros(0), d["jump_lengths"], sp.csr_matrix((0, j.shape[0])),
                        np.zeros(0), np.zeros(0), d["ncells"], None)
    oxj, osj = OC.solve(oj, rj, dt=d["dt"], density=d["density"], reference_length=1.0, tolerance=TOL)
    assert sj.converged and osj["converged"] and sj.pressure_unknowns == 0
    assert sj.iterations == osj["iterations"]
    assert rel(xj, oxj) <= 1e-10
    solver.close()
