"""CPU checks that pin the oracle's solve driver (stopping target, round-off
floor and its guard, restart-and-verify loop, nullspace convention) to the
reference's own known answers (no GPU).

Restated TwoPhaseAMR.jl tests:
  final iterate cannot inflate the stopping target    test/solvers.jl:1127-1182
  guarded iterate round-off floor (Float32, Float64)   test/solvers.jl:1184-1245
  round-off-limited heterogeneous Poisson solve        test/solvers.jl:1247-1277
  periodic nullspace convention                        test/operators/poisson.jl:921-933
The mock Krylov workspaces of test/solvers.jl:698-751 become scripted iterates
of the oracle's driver (`oracle.scripted_solve`), which is the same code the
oracle's mesh solves run through.
"""
import math

import numpy as np
import pytest

from oracle import OMesh, OSolver, assembled_floor, scripted_solve

# test/solvers.jl:1130-1133: a singular 2x2 Neumann pair
PAIR = np.array([[1.0, -1.0], [-1.0, 1.0]])
PAIR_V = np.ones(2)
PAIR_B = np.array([1.0, -1.0])


@pytest.mark.parametrize("relative_to_rhs", [False, True])
def test_final_iterate_cannot_inflate_the_target(relative_to_rhs):
    """An iterate with a huge nullspace part that has not reduced b - A x
    (test/solvers.jl:1127-1163): the floor of that iterate would exceed the
    residual, so it must not become the target."""
    initial_floor = assembled_floor(PAIR, PAIR_V, np.zeros(2), PAIR_B)
    x, st, runs = scripted_solve(PAIR, PAIR_V, PAIR_B, [[1.0e15, 1.0e15]], rtol=0.1, max_cycles=1,
                                 relative_to_rhs=relative_to_rhs)
    final_floor = assembled_floor(PAIR, PAIR_V, x, PAIR_B)
    actual = np.max(np.abs(PAIR_B - PAIR @ x))
    assert initial_floor < 0.1
    assert final_floor > actual
    assert np.array_equal(x, np.full(2, 1.0e15))
    assert st["cycles"] == 1 and runs == 1
    assert st["initial_linf"] == 1.0
    assert st["target_linf"] == 0.1
    assert st["final_linf"] == actual == 1.0
    assert actual > st["target_linf"]
    assert not st["converged"]


@pytest.mark.parametrize("relative_to_rhs", [False, True])
def test_partially_inflated_iterate_keeps_the_target(relative_to_rhs):
    """A large nullspace drift with a modest residual decrease
    (test/solvers.jl:1164-1181): the iterate's floor is not below
    initial / 100, so the target stays rtol * initial."""
    x, st, _ = scripted_solve(PAIR, PAIR_V, PAIR_B, [[1.0e15 + 0.5, 1.0e15]], rtol=0.1, max_cycles=1,
                              relative_to_rhs=relative_to_rhs)
    assert st["initial_linf"] == 1.0
    assert st["final_linf"] == 0.5
    assert st["target_linf"] == 0.1
    assert not st["converged"]


def _controlled(single, offset, corrections, relative_to_rhs, initial=None):
    """`controlled_solve` of test/solvers.jl:1185-1206: run k leaves
    [offset + corrections[k], offset]; rtol = 1e-4, max_cycles = len(corrections)."""
    its = [[offset + c, offset] for c in corrections]
    x, st, runs = scripted_solve(PAIR, PAIR_V, PAIR_B, its, rtol=1.0e-4, max_cycles=len(corrections),
                                 relative_to_rhs=relative_to_rhs, x0=initial, single=single)
    floor = assembled_floor(PAIR, PAIR_V, x, PAIR_B, single=single)
    return st, runs, floor


@pytest.mark.parametrize("single", [True, False], ids=["Float32", "Float64"])
@pytest.mark.parametrize("relative_to_rhs", [False, True])
def test_guarded_iterate_roundoff_floor(single, relative_to_rhs):
    """test/solvers.jl:1208-1244: the iterate's floor is accepted after
    progress when it is below initial / 100, rejected otherwise; a temporary
    increase at round-off does not block a later pass; a warm start already at
    the floor needs no iteration."""
    f = np.float32 if single else np.float64
    precision = 23 if single else 52
    small = math.ldexp(1.0, precision - 16)
    large = math.ldexp(1.0, precision - 12)
    corr = 511.0 / 512.0
    tiny = float(f(1.0e-4))

    st, runs, floor = _controlled(single, small, [corr], relative_to_rhs)
    assert runs == st["cycles"] == 1
    assert st["initial_linf"] == 1.0
    assert st["final_linf"] == 1.0 / 512.0
    assert floor < st["initial_linf"] / 100
    assert st["final_linf"] > tiny
    assert st["target_linf"] == floor
    assert st["converged"]

    st, runs, floor = _controlled(single, large, [corr], relative_to_rhs)
    assert runs == st["cycles"] == 1
    assert st["final_linf"] == 1.0 / 512.0
    assert floor > st["initial_linf"] / 100
    assert st["final_linf"] < floor
    assert st["target_linf"] == tiny
    assert not st["converged"]

    st, runs, floor = _controlled(single, small, [509 / 512, 508 / 512, 511 / 512], relative_to_rhs)
    assert runs == st["cycles"] == 3
    assert st["converged"]
    assert st["final_linf"] == 1.0 / 512.0
    assert st["target_linf"] == floor

    st, runs, floor = _controlled(single, small, [corr], relative_to_rhs, initial=[small + corr, small])
    assert runs == st["cycles"] == 0
    assert st["converged"]
    assert st["initial_linf"] == st["final_linf"] == 1.0 / 512.0
    assert st["target_linf"] == floor


def test_roundoff_limited_heterogeneous_poisson():
    """test/solvers.jl:1247-1277: 32^2 periodic cells in 4^2 blocks, face
    coefficient 1 left of x = 1/2 and 1e-3 right of it, the first cell's
    volume scaled by 1e-3 (a tiny cut cell), rhs sin(2 pi x) cos(2 pi y),
    rtol = 1e-15: the solve converges on the guarded floor of its own iterate."""
    m = OMesh(base_mesh=(32, 32), block_size=(4, 4))
    kappa = np.where(m.face_center[:, 0] < 0.5, 1.0, 1.0e-3)
    volumes = m.volumes.copy()
    volumes[0] *= 1.0e-3
    x, y = m.cell_centers()
    rhs = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    s = OSolver(m, kappa, volumes=volumes)
    initial_floor = s.floor(np.zeros(m.ncells), rhs * volumes, assembled=True)
    p, st = s.solve(rhs, rtol=1.0e-15, max_cycles=500)
    b = rhs * volumes - volumes * st["removed_rhs_mean"]
    final_floor = s.floor(p, b, assembled=True)
    assert st["converged"]
    assert st["final_linf"] <= st["target_linf"]
    assert st["final_linf"] > initial_floor
    assert st["target_linf"] > initial_floor
    assert math.isclose(st["target_linf"], final_floor, rel_tol=128 * np.finfo(float).eps)
    assert st["target_linf"] < st["initial_linf"] / 100
    assert st["final_linf"] < 1.0e-8
    # the factorized action's floor (the solver's own) is the assembled one
    # up to summation order on this uniform mesh
    assert math.isclose(s.floor(p, b), final_floor, rel_tol=1e-13)


def test_periodic_nullspace_convention():
    """test/operators/poisson.jl:921-933: a constant rhs on a periodic mesh is
    all nullspace: :error rejects it, :project removes mean 3 and returns the
    zero-mean pressure from an initial 7."""
    m = OMesh(block_size=(8, 8))
    s = OSolver(m, coarse_size=4)
    rhs = np.full(m.ncells, 3.0)
    with pytest.raises(RuntimeError, match="nonzero volume mean"):
        s.solve(rhs, x0=np.full(m.ncells, 7.0), compatibility="error")
    with pytest.raises(ValueError):
        s.solve(rhs, x0=np.full(m.ncells, 7.0), compatibility="invalid")
    p, st = s.solve(rhs, x0=np.full(m.ncells, 7.0), compatibility="project")
    eps = np.finfo(float).eps
    assert st["converged"]
    assert math.isclose(st["removed_rhs_mean"], 3.0, rel_tol=math.sqrt(eps))
    assert abs(np.dot(m.volumes, p)) <= 16 * eps
    assert np.max(np.abs(p)) <= 16 * eps
