"""Benchmark: Mcells/s to relative residual 1e-10 of the pressure solve.

Workload (default): unit square with walls, 16x16 blocks, uniform level 5,
refined to level 10 within 4 finest cells of 64 seeded circles; density ratio
1000:1 (mixture face coefficients from exact disk fractions); rhs = -div u of
a seeded smooth velocity. One step = one full solve from x = 0 on the GPU
with the rhs already resident in HBM (FP64 outer CG, the reference's strong
L-infinity stopping rule). The headline value uses the FP32 V-cycle; the
all-FP64 GPU solver runs on the same system for comparison.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--levels 5 10]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "twophaseamr.jl_amd"))

METRIC = "Mcells/s to rel. residual 1e-10, FP32-MG vs FP64 GPU solver; smoother HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_solves(solver, d_rhs, d_x, steps, warmup, rtol):
    import torch
    for _ in range(warmup):
        d_x.zero_()
        torch.cuda.synchronize()
        solver.solve_device(d_rhs.data_ptr(), d_x.data_ptr(), rtol=rtol, max_cycles=200)
    torch.cuda.synchronize()
    walls, dev = [], []
    st = None
    for _ in range(steps):
        d_x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solver.solve_device(d_rhs.data_ptr(), d_x.data_ptr(), rtol=rtol, max_cycles=200)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
        dev.append(st.solve_ms * 1e-3)
    return float(np.mean(walls)), float(np.mean(dev)), st


def cpu_baseline(levels, rtol):
    """The oracle (CPU FP64 restatement, one thread) on a bounded sample of the
    same workload family: the circles mesh at reduced maximum level."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import OMesh, OSolver
    import tpamr
    from tpamr import workloads as W
    mesh, circles = W.circles_mesh(levels)
    om = OMesh(base_mesh=mesh.base_mesh, block_size=mesh.block_size, periodic=mesh.periodic,
               max_level=mesh.max_level, refined_keys=mesh.refined_keys)
    c = W.liquid_fraction(mesh, circles)
    kappa = om.mixture_coefficients(c, 1000.0, 1.0)
    u = W.smooth_velocity(mesh)
    rhs = -om.divergence(u)
    t0 = time.perf_counter()
    s = OSolver(om, kappa)
    t1 = time.perf_counter()
    x, st = s.solve(rhs, rtol=rtol, max_cycles=200)
    t2 = time.perf_counter()
    return dict(value=mesh.ncells / (t2 - t1) / 1e6, unit="Mcells/s", cores=1, kind="port",
                sample=f"oracle FP64 solve (CPU restatement, 1 thread) on the circles workload at levels "
                       f"{levels[0]}-{levels[1]} ({mesh.ncells} cells, {st['cycles']} CG iterations, "
                       f"solve {t2 - t1:.2f} s, hierarchy setup {t1 - t0:.2f} s excluded)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--levels", type=int, nargs=2, default=(5, 10))
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-levels", type=int, nargs=2, default=(5, 7))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import tpamr
    from tpamr import workloads as W

    t0 = time.perf_counter()
    mesh, circles = W.circles_mesh(tuple(args.levels))
    c = W.liquid_fraction(mesh, circles)
    log(f"mesh {mesh} ({time.perf_counter() - t0:.1f} s)")
    u = W.smooth_velocity(mesh, device="cuda:0")
    log(f"velocity ({time.perf_counter() - t0:.1f} s)")

    results = {}
    x_sol = {}
    for prec in ("fp32_vcycle", "fp64"):
        t1 = time.perf_counter()
        solver = tpamr.EllipticSolver(tpamr.PoissonOperator(mesh, 1.0), precision=prec)
        tpamr.refresh_two_phase(solver, c, (1000.0, 1.0))
        log(f"{prec}: hierarchy {[lv['ncells'] for lv in solver.levels]} ({time.perf_counter() - t1:.1f} s)")
        rhs = -tpamr.divergence(solver, u)
        d_rhs = torch.as_tensor(rhs, device="cuda:0")
        d_x = torch.zeros(mesh.ncells, dtype=torch.float64, device="cuda:0")
        wall, dev, st = timed_solves(solver, d_rhs, d_x, args.steps, args.warmup, args.rtol)
        x_sol[prec] = d_x.cpu().numpy()
        res = dict(ms_per_step=wall * 1e3, device_ms=dev * 1e3, mcells=mesh.ncells / wall / 1e6,
                   cycles=st.cycles, converged=st.converged, final_linf=st.final_linf,
                   target_linf=st.target_linf, initial_linf=st.initial_linf)
        if prec == "fp32_vcycle":
            ms, nbytes = solver.time_smoother(reps=50)
            res["smoother_ms"] = ms
            res["smoother_bytes"] = nbytes
            res["levels"] = len(solver.levels)
        results[prec] = res
        log(f"{prec}: {res}")
        solver.close()
        del d_rhs, d_x
        torch.cuda.empty_cache()

    f32, f64 = results["fp32_vcycle"], results["fp64"]
    err = float(np.max(np.abs(x_sol["fp32_vcycle"] - x_sol["fp64"])) / np.max(np.abs(x_sol["fp64"])))
    achieved = f32["smoother_bytes"] / (f32["smoother_ms"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": f32["mcells"],
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 outer CG",
        "data": "synthetic",
        "config": {"workload": f"circles64_levels{args.levels[0]}-{args.levels[1]}_rho1000",
                   "cells": mesh.ncells, "blocks": mesh.nblocks, "block_size": list(mesh.block_size),
                   "mg_levels": f32["levels"], "rtol": args.rtol, "density_ratio": 1000.0},
        "cycles": f32["cycles"],
        "converged": bool(f32["converged"]),
        "fp64_solver": {"value": f64["mcells"], "ms_per_step": f64["ms_per_step"], "cycles": f64["cycles"],
                        "converged": bool(f64["converged"])},
        "speedup_vs_fp64": f32["mcells"] / f64["mcells"],
        "pressure_error_fp32_vs_fp64": err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_sweep<float,16,16,FINE>", "bytes_per_launch": f32["smoother_bytes"],
                     "launch_ms": f32["smoother_ms"]},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(tuple(args.cpu_levels), args.rtol)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
