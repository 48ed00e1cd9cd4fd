"""The oracle's CMDGFM block solve (oracle/oracle_cmdgfm.py) pinned by
closed forms, textbook routines and the reference's own Schur-complement
test:

- BiCGSTAB (the restatement of Krylov.jl 0.10's bicgstab! with a right
  preconditioner) takes the same iterates as SciPy's right-preconditioned
  BiCGSTAB, stops at the first iterate whose recursive residual is within
  rtol ||b||, and converges in one iteration with the exact inverse as
  preconditioner;
- the Crout ILU (IncompleteLU.jl 0.2's ilu) is the exact LU without dropping,
  compares L entries with tau before the division by the pivot, drops small
  U entries but keeps the pivot, and eliminates with the dropped factors;
- the jump Schur complement equals the dense J - C (B ./ diag(P)) of the
  reference's test (test/two_phase/two_velocity/projection/
  cmdgfm_projection.jl:344-355);
- the reduced-pressure block inverts the gauged pressure block exactly when
  its V-cycle is an exact solve (one and three components), and the block
  preconditioner with exact blocks solves a block-triangular system in one
  BiCGSTAB iteration;
- whole solves converge to the direct solution with every defect within the
  tolerance, also without jumps and from a previous increment.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle_cmdgfm as OC
from cmdgfm_cases import build_system
from oracle import OMesh, OSolver


def random_system(n, seed, density=0.15, dominance=1.5):
    rng = np.random.default_rng(seed)
    a = sp.random(n, n, density=density, random_state=rng, format="csr")
    a.data = rng.uniform(-1.0, 1.0, a.nnz)
    d = np.asarray(abs(a).sum(axis=1)).ravel() * dominance + 1.0
    return sp.csr_matrix(a + sp.diags(d))


def scipy_iterate(a, b, k, m=None):
    x, _ = sla.bicgstab(a, b, rtol=1e-300, atol=0.0, maxiter=k, M=m)
    return x


@pytest.mark.parametrize("precondition", [False, True])
def test_bicgstab_iterates_equal_scipy(precondition):
    a = random_system(60, 1)
    b = np.random.default_rng(2).standard_normal(60)
    dinv = 1.0 / a.diagonal()
    m = sla.LinearOperator(a.shape, matvec=lambda v: dinv * v) if precondition else None
    pc = (lambda v: dinv * v) if precondition else None
    for k in range(1, 7):
        x, its, _ = OC.bicgstab(lambda v: a @ v, b, pc, atol=0.0, rtol=0.0, itmax=k)
        assert its == k
        ref = scipy_iterate(a, b, k, m)
        assert np.max(np.abs(x - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_bicgstab_stopping_rule_and_convergence():
    a = random_system(80, 3, dominance=1.1)
    b = np.random.default_rng(4).standard_normal(80)
    rtol = 1e-9
    x, its, status = OC.bicgstab(lambda v: a @ v, b, None, atol=0.0, rtol=rtol)
    assert status == "solved"
    # the recursive residual tracks the true one; the stop is the first
    # iterate with ||r|| <= rtol ||b||
    true = np.linalg.norm(b - a @ x)
    assert true <= 1.01 * rtol * np.linalg.norm(b)
    prev = scipy_iterate(a, b, its - 1)
    assert np.linalg.norm(b - a @ prev) > rtol * np.linalg.norm(b)
    assert np.max(np.abs(x - sla.spsolve(a.tocsc(), b))) <= 1e-7 * np.max(np.abs(x))
    # itmax = 0 means 2n; a zero rhs is solved by zero in no iteration
    x0, its0, _ = OC.bicgstab(lambda v: a @ v, np.zeros(80))
    assert its0 == 0 and not np.any(x0)


def test_bicgstab_exact_preconditioner_one_iteration():
    a = random_system(40, 5)
    b = np.random.default_rng(6).standard_normal(40)
    inv = np.linalg.inv(a.toarray())
    x, its, status = OC.bicgstab(lambda v: a @ v, b, lambda v: inv @ v, atol=0.0, rtol=1e-12)
    assert its == 1 and status == "solved"
    assert np.max(np.abs(a @ x - b)) <= 1e-12 * np.max(np.abs(b))


def test_crout_ilu_without_dropping_is_exact_lu():
    s = random_system(50, 7, density=0.08)
    lo, up = OC.crout_ilu(s, 0.0)
    assert sp.triu(lo).nnz == 0
    assert sp.tril(up, -1).nnz == 0
    prod = (sp.eye(50) + lo) @ up
    assert np.max(np.abs((prod - s).toarray())) <= 1e-13 * np.max(np.abs(s.toarray()))
    # the ILU solve is then the exact solve
    b = np.random.default_rng(8).standard_normal(50)
    assert np.max(np.abs(s @ OC.ilu_solve(lo, up, b) - b)) <= 1e-12 * np.max(np.abs(b))


def test_crout_ilu_drop_rule():
    tau = 0.1
    # L entry 0.5 (unscaled) >= tau is kept as 0.5 / 100; a scaled
    # comparison would drop it. U entry 0.05 < tau is dropped; the pivot 0.01
    # is kept although below tau.
    s = sp.csr_matrix(np.array([[100.0, 0.05, 0.0], [0.5, 0.01, 0.0], [0.0, 0.0, 1.0]]))
    lo, up = OC.crout_ilu(s, tau)
    assert lo[1, 0] == 0.5 / 100.0
    assert up[0, 1] == 0.0 and up.nnz == 3
    assert up[1, 1] == 0.01
    # elimination with the stored factors: u_11 = 0.01 - l_10 * (dropped
    # u_01) = 0.01, not 0.01 - 0.005 * 0.05
    dense = np.array([[4.0, 2.0, 0.05], [1.0, 3.0, 1.0], [0.02, 1.0, 5.0]])
    lo, up = OC.crout_ilu(sp.csr_matrix(dense), tau)
    l10 = 1.0 / 4.0
    assert lo[1, 0] == l10
    assert up[0, 2] == 0.0  # 0.05 < tau
    u12 = 1.0 - l10 * 0.0   # from the stored (dropped) u_02
    assert up[1, 2] == u12
    assert lo[2, 0] == 0.0  # 0.02 < tau (before scaling)
    u11 = 3.0 - l10 * 2.0
    l21 = (1.0 - 0.0 * 2.0) / u11
    assert lo[2, 1] == l21
    assert up[2, 2] == 5.0 - (0.0 * 0.05 + l21 * u12)


def _pressure_case(kw, kappa_fn=None):
    om = OMesh(**kw)
    cx, cy = om.cell_centers()
    m, p = om.face_minus, om.face_plus
    if kappa_fn is None:
        rho = np.where(np.hypot(cx - 0.5, cy - 0.5) < 0.27, 1000.0, 1.0)
        kap = 2.0 / (rho[m] + rho[p])
    else:
        kap = kappa_fn(om)
    return om, OSolver(om, kap, coarse_size=16)


def _strips(om):
    k = np.ones(om.nfaces)
    k[(om.face_axis == 0) & np.isin(om.face_center[:, 0], [0.25, 0.75])] = 0.0
    return k


def _block(d, vcycle):
    return OC.BlockSystem(d["matrix"], d["np"], d["pressure_indices"], d["pressure_anchors"], d["pressure_volumes"],
                          d["jump_lengths"], d["omitted"], d["omitted_rhs"], d["omitted_volumes"], d["ncells"],
                          vcycle)


def test_jump_schur_matches_reference_dense_formula():
    om, osol = _pressure_case(dict(base_mesh=(32, 32), block_size=(8, 8), periodic=(False, False)))
    d = build_system(om, osol)
    a = d["matrix"].toarray()
    npp = d["np"]
    pres, b, c, j = a[:npp, :npp], a[:npp, npp:], a[npp:, :npp], a[npp:, npp:]
    expected = j - c @ (b / np.diag(pres)[:, None])
    got = OC.jump_schur(d["matrix"], npp).toarray()
    assert np.allclose(got, expected, rtol=1e-13, atol=1e-13 * np.max(np.abs(expected)))
    assert not np.array_equal(got, j)


@pytest.mark.parametrize("strips", [False, True])
def test_reduced_pressure_block_with_an_exact_vcycle(strips):
    kw = dict(base_mesh=(16, 16), block_size=(8, 8), periodic=(False, False))
    om, osol = _pressure_case(kw, _strips if strips else None)
    d = build_system(om, osol, jumps=False)
    s = osol.level_matrix(0).toarray()
    pinv = np.linalg.pinv(s)

    def exact(r):
        return pinv @ r

    bs = _block(d, exact)
    assert d["components"] == (3 if strips else 1)
    rhs = np.random.default_rng(9).standard_normal(d["np"])
    y = bs.reduced_pressure_ldiv(rhs)
    a_pp = d["matrix"].toarray()
    assert np.max(np.abs(a_pp @ y - rhs)) <= 1e-9 * np.max(np.abs(rhs))


def test_block_preconditioner_with_exact_blocks_solves_in_one_iteration():
    kw = dict(base_mesh=(16, 16), block_size=(8, 8), periodic=(False, False))
    om, osol = _pressure_case(kw)
    d = build_system(om, osol)
    npp = d["np"]
    a = d["matrix"].tolil()
    a[:npp, npp:] = 0.0   # B = 0: block lower triangular, Schur = J
    a = sp.csr_matrix(a)
    a.eliminate_zeros()
    pinv = np.linalg.pinv(osol.level_matrix(0).toarray())
    bs = OC.BlockSystem(a, npp, d["pressure_indices"], d["pressure_anchors"], d["pressure_volumes"],
                        d["jump_lengths"], sp.csr_matrix((0, d["matrix"].shape[0])), np.zeros(0), np.zeros(0),
                        d["ncells"], lambda r: pinv @ r, drop_tolerance=0.0)
    b = d["rhs"]
    x, its, status = OC.bicgstab(lambda v: a @ v, b, bs.block_ldiv, atol=0.0, rtol=1e-10)
    assert its == 1 and status == "solved"


CASES = {
    "uniform32": dict(base_mesh=(32, 32), block_size=(8, 8), periodic=(False, False)),
    "periodic32": dict(base_mesh=(32, 32), block_size=(8, 8), periodic=(True, True)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_solve_converges_to_the_direct_solution(name):
    om, osol = _pressure_case(CASES[name])
    d = build_system(om, osol)
    bs = _block(d, osol.vcycle)
    tol = 1e-8
    x, st = OC.solve(bs, d["rhs"], dt=d["dt"], density=d["density"], reference_length=1.0, tolerance=tol)
    assert st["converged"] and st["jump_unknowns"] > 0
    assert st["pressure_defect"] <= tol and st["jump_defect"] <= tol and st["anchor_defect"] <= tol
    xd = sla.spsolve(d["matrix"].tocsc(), d["rhs"])
    assert np.max(np.abs(x - xd)) <= 1e-6 * np.max(np.abs(xd))
    # with the true solution's pressures as the previous increment the
    # warm start takes them (zero at the anchors) and little is left to do
    prev = np.zeros(d["ncells"])
    prev[d["pressure_indices"]] = xd[:d["np"]]
    x2, st2 = OC.solve(bs, d["rhs"], dt=d["dt"], density=d["density"], reference_length=1.0, tolerance=tol,
                       previous_increment=prev)
    assert st2["initial_guess_used"] and st2["converged"]
    assert st2["iterations"] < st["iterations"]


def test_solve_without_jumps():
    om, osol = _pressure_case(dict(base_mesh=(16, 16), block_size=(8, 8), periodic=(False, False)), _strips)
    d = build_system(om, osol, jumps=False)
    bs = _block(d, osol.vcycle)
    x, st = OC.solve(bs, d["rhs"], dt=d["dt"], density=d["density"], reference_length=1.0, tolerance=1e-8)
    assert st["converged"] and st["jump_unknowns"] == 0 and not st["initial_guess_used"]
    assert st["anchor_defect"] <= 1e-8
    xd = sla.spsolve(d["matrix"].tocsc(), d["rhs"])
    assert np.max(np.abs(x - xd)) <= 1e-6 * np.max(np.abs(xd))
