"""Adjoint of forcing-plus-projection steps and of a checkpointed run of them
(tpamr_project_adjoint, tpamr.ProjectionRun). These have no reference
counterpart; they are pinned by
- transposition identities of one step: <ubar, P v> = <P^T ubar, v> and
  <ubar, -dt G A^+ V d> = <dJ/db, d>, exact up to the solve tolerance;
- central finite differences of J along seeded directions (the first step's
  pressure rhs, the per-step face coefficients, the initial velocity);
- the same gradient composed from the oracle's FP64 CPU primitives
  (project, weighted_jump_force, divergence, solve, the inverse metric K).
The run is the Stokes-wave-like two-phase case at a small size: [0, 2] x [0, 1],
8x8 blocks, levels 1-4, density (1000, 1.2), gravity (0, -9.81), ten steps of
gravity! plus project! with the interface at each step's time, J = kinetic
energy after the last step."""
import numpy as np
import pytest

import tpamr
from tpamr import workloads as W
from oracle import OMesh, OSolver

pytestmark = pytest.mark.gpu

DT = 1e-3
STEPS = 10


def wave(levels=(1, 4), block=8):
    pm = W.wave_mesh(levels, block=block)
    om = OMesh(base_mesh=pm.base_mesh, block_size=pm.block_size, extent=pm.extent, periodic=pm.periodic,
               max_level=pm.max_level, refined_keys=pm.refined_keys)
    return pm, om


def jump(pm, x):
    return pm.face_area * (x[pm.face_plus] - x[pm.face_minus])


def volumes_mean_free(pm, db):
    v = pm.volumes
    return db - v * (np.sum(db) / np.sum(v))


@pytest.fixture(scope="module")
def case():
    pm, om = wave()
    gs = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, np.ones(pm.nfaces)), precision="fp64", device=0)
    steps, frac_end = W.wave_steps(gs, STEPS, DT)
    u0 = 0.01 * W.smooth_velocity(pm)
    yield pm, om, gs, steps, frac_end, u0
    gs.close()


def test_step_adjoint_transpose_identities(case):
    pm, om, gs, steps, _, _ = case
    kappa, jumps = steps[3]
    gs.operator.replace_face_coefficient(kappa)
    gs.refresh()
    rng = np.random.default_rng(21)
    v = rng.standard_normal(pm.nfaces)
    ub = rng.standard_normal(pm.nfaces)
    u = v.copy()
    _, phi = tpamr.project(u, np.zeros(pm.ncells), gs, DT, tolerance=1e-16)
    us, db, dk, st = tpamr.project_adjoint(gs, ub, DT, phi=phi, rtol=1e-13)
    assert st.converged
    lhs, rhs = ub @ u, us @ v
    assert abs(lhs - rhs) <= 1e-10 * (np.abs(ub) @ np.abs(u) + np.abs(us) @ np.abs(v))
    d = rng.standard_normal(pm.ncells)
    q = np.zeros(pm.ncells)
    gs.solve(q, d, rtol=1e-13)
    lhs = -DT * (ub @ tpamr.weighted_gradient(gs, q))
    assert abs(lhs - db @ d) <= 1e-9 * np.abs(db) @ np.abs(d)
    # the gauge of lambda: dJ/db is volume-mean free
    assert abs(np.sum(db)) <= 1e-10 * np.sum(np.abs(db))


def test_step_adjoint_kappa_gradient_matches_finite_differences(case):
    pm, om, gs, steps, _, _ = case
    kappa, jumps = steps[5]
    rng = np.random.default_rng(22)
    v = rng.standard_normal(pm.nfaces)
    ub = rng.standard_normal(pm.nfaces)

    def J(k):
        gs.operator.replace_face_coefficient(k)
        gs.refresh()
        u = v + DT * tpamr.weighted_jump_force(gs, jumps)
        _, phi = tpamr.project(u, np.zeros(pm.ncells), gs, DT, tolerance=1e-16)
        return ub @ u, phi

    _, phi = J(kappa)
    _, _, dk, _ = tpamr.project_adjoint(gs, ub, DT, phi=phi, jumps=jumps, rtol=1e-13)
    for s in range(3):
        d = np.random.default_rng(100 + s).standard_normal(pm.nfaces)
        eps = 1e-5
        jp, _ = J(kappa * (1 + eps * d))
        jm, _ = J(kappa * (1 - eps * d))
        fd = (jp - jm) / (2 * eps)
        ad = dk @ (kappa * d)
        assert abs(fd - ad) <= 1e-6 * max(abs(fd), 1e-30), (s, fd, ad)


def oracle_run(om, pm, steps, u0, frac_end):
    """Forward and reverse sweep of the run with the oracle's FP64 CPU
    primitives, one oracle solver per step's coefficients."""
    K, _ = om.metric()
    u = u0.copy()
    solvers, phis = [], []
    for kappa, jumps in steps:
        s = OSolver(om, kappa)
        u = u + DT * s.weighted_jump_force(jumps)
        u, _, phi, st = s.project(u, np.zeros(om.ncells), DT, tolerance=1e-16)
        solvers.append(s)
        phis.append(phi)
    J, ub = tpamr.kinetic_energy(pm, u, frac_end, W.WAVE_DENSITY)
    gk = [None] * len(steps)
    for n in range(len(steps) - 1, -1, -1):
        s, (kappa, jumps) = solvers[n], steps[n]
        rhs = -om.divergence(s.weighted_jump_force(ub))
        lam, st = s.solve(rhs, rtol=1e-13)
        us = ub - jump(pm, lam)
        sk = np.sqrt(kappa)
        Wv, Z = sk * us, sk * (jumps - jump(pm, phis[n]))
        gk[n] = DT * (Wv * (K @ Z) + Z * (K @ Wv)) / (2 * kappa)
        db = -DT * pm.volumes * lam
        ub = us
    return J, volumes_mean_free(pm, db), gk, ub


def test_run_gradient_matches_finite_differences_and_oracle(case):
    pm, om, gs, steps, frac_end, u0 = case
    run = tpamr.ProjectionRun(gs, DT, tolerance=1e-16, adjoint_rtol=1e-13)
    uN = run.forward(u0, steps)
    assert len(run.checkpoints) == STEPS and all(st.converged for st in run.stats)
    J, ub = tpamr.kinetic_energy(pm, uN, frac_end, W.WAVE_DENSITY)
    g = run.backward(ub)
    db = g["rhs"]

    def Jof(**kw):
        u = run.forward(u0 if "u0" not in kw else kw["u0"], kw.get("steps", steps), rhs_shift=kw.get("shift"))
        return tpamr.kinetic_energy(pm, u, frac_end, W.WAVE_DENSITY)[0]

    rng = np.random.default_rng(23)
    b0 = tpamr.divergence(gs, u0) / DT
    for s in range(4):
        d = rng.standard_normal(pm.ncells)
        eps = 1e-2 * np.max(np.abs(b0)) / np.max(np.abs(d))  # J is quadratic in b_0: central differences are exact
        fd = (Jof(shift=eps * d) - Jof(shift=-eps * d)) / (2 * eps)
        assert abs(fd - db @ d) <= 1e-7 * (np.abs(db) @ np.abs(d)), (s, fd, db @ d)
    for n in (0, 4, 9):
        kappa, jumps = steps[n]
        d = rng.standard_normal(pm.nfaces)
        eps = 1e-5
        sp = list(steps)
        sp[n] = (kappa * (1 + eps * d), jumps)
        jp = Jof(steps=sp)
        sp[n] = (kappa * (1 - eps * d), jumps)
        jm = Jof(steps=sp)
        fd = (jp - jm) / (2 * eps)
        ad = g["kappa"][n] @ (kappa * d)
        assert abs(fd - ad) <= 1e-6 * abs(fd), (n, fd, ad)
    d = rng.standard_normal(pm.nfaces) * 1e-3
    fd = (Jof(u0=u0 + d) - Jof(u0=u0 - d)) / 2
    assert abs(fd - g["u0"] @ d) <= 1e-7 * abs(fd)
    # the FP64 CPU reference gradient
    Jo, dbo, gko, u0o = oracle_run(om, pm, steps, u0, frac_end)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert np.max(np.abs(db - dbo)) <= 1e-8 * np.max(np.abs(dbo))
    for n in range(STEPS):
        assert np.max(np.abs(g["kappa"][n] - gko[n])) <= 1e-8 * np.max(np.abs(gko[n])), n
    assert np.max(np.abs(g["u0"] - u0o)) <= 1e-8 * np.max(np.abs(u0o))


@pytest.mark.parametrize("precision", ["fp32_refine", "fp32_vcycle"])
def test_run_gradient_in_mixed_precision(case, precision):
    """The FP32-preconditioned modes reach the same FP64 stopping rule, so the
    checkpoints and the gradient agree with the FP64 run."""
    pm, om, gs, steps, frac_end, u0 = case
    ref = tpamr.ProjectionRun(gs, DT, tolerance=1e-16, adjoint_rtol=1e-13)
    J64, ub = tpamr.kinetic_energy(pm, ref.forward(u0, steps), frac_end, W.WAVE_DENSITY)
    g64 = ref.backward(ub)
    ms = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, steps[0][0]), precision=precision, device=0)
    run = tpamr.ProjectionRun(ms, DT, tolerance=1e-16, adjoint_rtol=1e-13)
    J, ub = tpamr.kinetic_energy(pm, run.forward(u0, steps), frac_end, W.WAVE_DENSITY)
    g = run.backward(ub)
    ms.close()
    assert abs(J - J64) <= 1e-10 * abs(J64)
    for n in range(STEPS):
        assert np.max(np.abs(run.checkpoints[n] - ref.checkpoints[n])) <= 1e-8 * np.max(np.abs(ref.checkpoints[n]))
    assert np.max(np.abs(g["rhs"] - g64["rhs"])) <= 1e-8 * np.max(np.abs(g64["rhs"]))
    assert np.max(np.abs(g["kappa"][0] - g64["kappa"][0])) <= 1e-8 * np.max(np.abs(g64["kappa"][0]))


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_device_run_matches_host_run(case, precision):
    """The run kept in HBM (device-memory entry points: coefficient refresh,
    force, projection, pullback) is the host-array run, bit for bit. Each run
    gets a fresh solver: the dense coarse inverse is updated from the one of
    the previous refresh, so two solvers agree bit for bit when they have seen
    the same sequence of coefficients."""
    import torch
    pm, om, gs, steps, frac_end, u0 = case

    def fresh():
        return tpamr.EllipticSolver(tpamr.PoissonOperator(pm, steps[0][0]), precision=precision, device=0)

    for rtol in (None, 1e-11):
        s = fresh()
        host = tpamr.ProjectionRun(s, DT, tolerance=1e-14, rtol=rtol, adjoint_rtol=1e-13)
        uh = host.forward(u0, steps)
        _, ubh = tpamr.kinetic_energy(pm, uh, frac_end, W.WAVE_DENSITY)
        gh = host.backward(ubh)
        s.close()
        s = fresh()
        dev = torch.device("cuda:0")
        dsteps = [(torch.as_tensor(k, device=dev), torch.as_tensor(j, device=dev)) for k, j in steps]
        run = tpamr.DeviceProjectionRun(s, DT, tolerance=1e-14, rtol=rtol, adjoint_rtol=1e-13)
        ud = run.forward(torch.as_tensor(u0, device=dev), dsteps)
        assert np.array_equal(ud.cpu().numpy(), uh)
        for a, b in zip(run.checkpoints, host.checkpoints):
            assert np.array_equal(a.cpu().numpy(), b)
        gd = run.backward(torch.as_tensor(ubh, device=dev))
        assert np.array_equal(gd["rhs"].cpu().numpy(), gh["rhs"])
        assert np.array_equal(gd["u0"].cpu().numpy(), gh["u0"])
        for a, b in zip(gd["kappa"], gh["kappa"]):
            assert np.array_equal(a.cpu().numpy(), b)
        s.close()


def test_device_run_from_interface_data_matches_run_from_steps(case):
    """DeviceProjectionRun.forward_interface computes each step's mixture
    coefficients and gravity jumps on the GPU from the step's interface data;
    the run equals the one from host-built steps (wave_steps), bit for bit."""
    import torch
    pm, om, gs, steps, frac_end, u0 = case
    dev = torch.device("cuda:0")
    ifaces = []
    for n, data in enumerate(W.wave_interfaces(pm, STEPS, DT)):
        if n == STEPS:
            assert np.array_equal(data[0], frac_end)
            break
        ifaces.append(tuple(torch.as_tensor(a, device=dev) for a in data))
    outs = []
    for how in ("steps", "interface"):
        s = tpamr.EllipticSolver(tpamr.PoissonOperator(pm, steps[0][0]), precision="fp64", device=0)
        run = tpamr.DeviceProjectionRun(s, DT, tolerance=1e-14, adjoint_rtol=1e-13)
        if how == "steps":
            uN = run.forward(torch.as_tensor(u0, device=dev), [(torch.as_tensor(k, device=dev),
                                                                 torch.as_tensor(j, device=dev)) for k, j in steps])
        else:
            uN = run.forward_interface(torch.as_tensor(u0, device=dev), ifaces, density=W.WAVE_DENSITY,
                                       gravity=W.WAVE_GRAVITY)
            for (k, j), (kh, jh) in zip(run.steps, steps):
                assert np.array_equal(k.cpu().numpy(), kh) and np.array_equal(j.cpu().numpy(), jh)
        g = run.backward(uN.clone())
        outs.append((uN.cpu().numpy(), [c.cpu().numpy() for c in run.checkpoints], g["rhs"].cpu().numpy()))
        s.close()
    (ua, ca, ga), (ub, cb, gb) = outs
    assert np.array_equal(ua, ub) and np.array_equal(ga, gb)
    assert all(np.array_equal(a, b) for a, b in zip(ca, cb))
