"""Benchmark: Mcells/s to relative residual 1e-10 of the pressure solve.

Workload (default): unit square with walls, 16x16 blocks, uniform level 5,
refined to level 10 within 4 finest cells of 64 seeded circles; density ratio
1000:1 (mixture face coefficients from exact disk fractions); rhs = -div u of
a seeded smooth velocity. One step = one full solve from x = 0 on the GPU
with the rhs already resident in HBM to the reference's strong L-infinity
stopping rule. The headline value is the 'fp32_refine' mode (FP64 iterative
refinement around an inner CG with FP32 vectors and FP64 inner products,
preconditioned by the FP32 V-cycle); the 'fp32_vcycle' mode (FP64 CG around
the FP32 V-cycle) and the all-FP64 GPU solver run on the same system for
comparison.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--levels 5 10]

Roofline of the dominant kernel (the finest-level FP32 Jacobi sweep): its
algorithmic bytes per launch are 24 B per cell (x, b, D^-1, two face
coefficients read, x' written, 4 B each) plus 12 B per stored entry of the
rows next to coarse-fine interfaces; the average launch time comes from HIP
events around launches of that kernel on the solver's stream.

Hardware-counter traffic (`roofline.traffic`) comes from two rocprofv3 PMC
passes over `--probe` (FETCH_SIZE, then WRITE_SIZE), calibrated against a
plain device copy of known size in the same run:

  rocprofv3 --pmc FETCH_SIZE -d <dir_f> -o pmc -- python3 bench.py --probe
  rocprofv3 --pmc WRITE_SIZE -d <dir_w> -o pmc -- python3 bench.py --probe
  python3 bench.py --pmc-summary <dir_f>/pmc_results.db <dir_w>/pmc_results.db

The summary is written to profiles/ and read back by later bench runs.
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
SWEEP_KERNEL = "k_sweep<float, 16, 16, true, false, double>"
COPY_BYTES = 1 << 28  # per direction, calibration copy of the PMC probe
TRAFFIC_MD = os.path.join(ROOT, "profiles", "k_sweep_fp32_circles64_levels5-10_traffic.md")
# precision modes solved on the workload; the headline value is the first
MODES = ("fp32_refine", "fp32_vcycle", "fp64")
HEADLINE = MODES[0]


def sweep_algorithmic_bytes(ncells, irregular_nnz, elem=4):
    """Algorithmic bytes of one finest-level Jacobi sweep: 6 words per cell
    (x, b, D^-1, two face coefficients, x') plus 12 B per irregular entry."""
    return 6 * elem * ncells + 12 * irregular_nnz


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


def host_cpu():
    """(model name, threads this process may run on; at most 16, the GPU
    box's CPU share per GPU)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return model, max(1, min(16, n))


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


def cpu_baseline(levels, rtol, steps=3):
    """The oracle (CPU FP64 restatement) on a bounded sample of the same
    workload family: the circles mesh at a reduced maximum level, solved on
    the host's threads (OpenMP; results independent of the thread count) and
    again on one thread. The same system is solved on the GPU in every
    precision mode and the pressures are compared with the CPU FP64 solution."""
    import torch
    O = _oracle()
    import tpamr
    from tpamr import workloads as W
    model, threads = host_cpu()
    mesh, circles = W.circles_mesh(levels)
    om = O.OMesh(base_mesh=mesh.base_mesh, block_size=mesh.block_size, periodic=mesh.periodic,
                 max_level=mesh.max_level, refined_keys=mesh.refined_keys)
    c = W.liquid_fraction(mesh, circles)
    kappa = om.mixture_coefficients(c, 1000.0, 1.0)
    u = W.smooth_velocity(mesh, device="cuda:0")
    rhs = -om.divergence(u)
    O.set_threads(threads)
    t0 = time.perf_counter()
    s = O.OSolver(om, kappa)
    t1 = time.perf_counter()
    x_cpu, st = s.solve(rhs, rtol=rtol, max_cycles=200)
    t2 = time.perf_counter()
    O.set_threads(1)
    x_one, st_one = s.solve(rhs, rtol=rtol, max_cycles=200)
    t3 = time.perf_counter()
    O.set_threads(threads)
    del s
    log(f"cpu baseline: {mesh.ncells} cells, setup {t1 - t0:.1f} s, solve {t2 - t1:.1f} s on {threads} threads, "
        f"{t3 - t2:.1f} s on 1")
    gpu = {"cells": mesh.ncells}
    d_rhs = torch.as_tensor(rhs, device="cuda:0")
    for prec in MODES:
        solver = tpamr.EllipticSolver(tpamr.PoissonOperator(mesh, kappa), precision=prec)
        d_x = torch.zeros(mesh.ncells, dtype=torch.float64, device="cuda:0")
        wall, _, gst = timed_solves(solver, d_rhs, d_x, steps, 1, rtol)
        x = d_x.cpu().numpy()
        gpu[prec] = {"mcells": mesh.ncells / wall / 1e6, "cycles": gst.cycles, "converged": bool(gst.converged),
                     "pressure_error_vs_cpu_fp64": float(np.max(np.abs(x - x_cpu)) / np.max(np.abs(x_cpu)))}
        solver.close()
    gpu["speedup_fp32_refine_vs_fp64"] = gpu["fp32_refine"]["mcells"] / gpu["fp64"]["mcells"]
    cpu = dict(value=mesh.ncells / (t2 - t1) / 1e6, unit="Mcells/s", cores=threads, kind="port",
               cpu_model=model,
               sample=f"oracle FP64 solve (CPU restatement, OpenMP on {threads} threads) on the circles workload "
                      f"at levels {levels[0]}-{levels[1]} ({mesh.ncells} cells, {st['cycles']} CG iterations, "
                      f"solve {t2 - t1:.2f} s, hierarchy setup {t1 - t0:.2f} s excluded)",
               cpu_cycles=st["cycles"],
               value_1_thread=mesh.ncells / (t3 - t2) / 1e6,
               thread_count_independent=bool(np.array_equal(x_cpu, x_one) and st["cycles"] == st_one["cycles"]))
    return cpu, gpu


def cpu_full(case, x_sol, rtol):
    """The oracle's FP64 CPU solve of the benchmark system itself (the full
    mesh, on the host's threads), and each GPU mode's pressure error against
    it."""
    O = _oracle()
    model, threads = host_cpu()
    O.set_threads(threads)
    mesh = case.mesh
    t0 = time.perf_counter()
    om = O.OMesh(base_mesh=mesh.base_mesh, block_size=mesh.block_size, periodic=mesh.periodic,
                 max_level=mesh.max_level, refined_keys=mesh.refined_keys)
    kappa = om.mixture_coefficients(case.fraction, 1000.0, 1.0)
    s = O.OSolver(om, kappa)
    t1 = time.perf_counter()
    log(f"cpu full: setup {t1 - t0:.1f} s")
    rhs = case.d_rhs.cpu().numpy()
    x_cpu, st = s.solve(rhs, rtol=rtol, max_cycles=200)
    t2 = time.perf_counter()
    log(f"cpu full: solve {t2 - t1:.1f} s, {st['cycles']} iterations")
    ref = np.max(np.abs(x_cpu))
    return dict(cells=mesh.ncells, threads=threads, cpu_model=model, setup_s=t1 - t0, solve_s=t2 - t1,
                mcells=mesh.ncells / (t2 - t1) / 1e6, cycles=st["cycles"], converged=bool(st["converged"]),
                target_linf=st["target_linf"], final_linf=st["final_linf"],
                pressure_error_vs_cpu_fp64={p: float(np.max(np.abs(x - x_cpu)) / ref) for p, x in x_sol.items()})


class Case:
    """Mesh, liquid fraction and device rhs of the circles workload; solvers
    of any precision mode on it."""

    def __init__(self, levels, device="cuda:0"):
        import torch
        from tpamr import workloads as W
        self.device = device
        self.mesh, circles = W.circles_mesh(tuple(levels))
        self.fraction = W.liquid_fraction(self.mesh, circles)
        self.u = W.smooth_velocity(self.mesh, device=device)
        self.d_rhs = None
        self._torch = torch

    def solver(self, precision):
        import tpamr
        s = tpamr.EllipticSolver(tpamr.PoissonOperator(self.mesh, 1.0), precision=precision)
        tpamr.refresh_two_phase(s, self.fraction, (1000.0, 1.0))
        if self.d_rhs is None:
            self.d_rhs = self._torch.as_tensor(-tpamr.divergence(s, self.u), device=self.device)
        return s


def probe(levels):
    """Short run for hardware-counter passes: one headline-mode solve, 10
    finest-level sweeps, then 5 device copies of COPY_BYTES (calibration; the
    last 5 dispatches of the process)."""
    import torch
    case = Case(levels)
    mesh, solver = case.mesh, case.solver(HEADLINE)
    d_rhs = case.d_rhs
    d_x = torch.zeros(mesh.ncells, dtype=torch.float64, device="cuda:0")
    st = solver.solve_device(d_rhs.data_ptr(), d_x.data_ptr(), rtol=1e-10, max_cycles=200)
    log(f"probe: {mesh.ncells} cells, solve {st.cycles} iterations")
    solver.time_smoother(reps=10)
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    info = solver.level_info(0)
    print(json.dumps({"probe": True, "cells": mesh.ncells, "irregular_nnz": info["irregular_nnz"]}), flush=True)


def pmc_summary(db_fetch, db_write, out_path=TRAFFIC_MD):
    """Per-launch HBM traffic of the finest-level FP32 sweep from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch), each scaled by what the counter reads for
    the calibration copy (COPY_BYTES read and written)."""
    import sqlite3

    def rows(db, counter):
        con = sqlite3.connect(db)
        r = con.execute("select dispatch_id, kernel_name, grid_size, value from counters_collection "
                        "where counter_name = ? order by dispatch_id", (counter,)).fetchall()
        con.close()
        return r

    out = {}
    for counter, db in (("FETCH_SIZE", db_fetch), ("WRITE_SIZE", db_write)):
        r = rows(db, counter)
        copy = r[-5:]
        sweeps = [x for x in r if SWEEP_KERNEL in x[1]]
        gmax = max(x[2] for x in sweeps)
        sweeps = [x[3] for x in sweeps if x[2] == gmax]
        copy_kib = float(np.mean([x[3] for x in copy]))
        factor = (COPY_BYTES / 1024.0) / copy_kib
        out[counter] = dict(sweep_kib=float(np.mean(sweeps)), n=len(sweeps), copy_kib=copy_kib, factor=factor,
                            bytes=float(np.mean(sweeps)) * 1024.0 * factor)
    f, w = out["FETCH_SIZE"], out["WRITE_SIZE"]
    total = f["bytes"] + w["bytes"]
    lines = [
        "# HBM traffic of the finest-level FP32 Jacobi sweep, circles64 levels 5-10",
        "",
        "Commands (one MI355X; `cd /tmp && export TMPDIR=/tmp` first; `<repo>` the repository root, `<out>` an "
        "output folder):",
        "",
        "```",
        "rocprofv3 --pmc FETCH_SIZE -d <out>/pmc_f -o pmc -- python3 <repo>/bench.py --probe",
        "rocprofv3 --pmc WRITE_SIZE -d <out>/pmc_w -o pmc -- python3 <repo>/bench.py --probe",
        "python3 bench.py --pmc-summary <out>/pmc_f/pmc_results.db <out>/pmc_w/pmc_results.db",
        "```",
        "",
        f"Kernel `{SWEEP_KERNEL}` at the finest level ({f['n']} dispatches per pass). Calibration: the last 5",
        f"dispatches of the probe are a device copy reading and writing {COPY_BYTES} B each.",
        "",
        "| counter | sweep KiB/dispatch | copy KiB/dispatch | scale (known/counted) | sweep bytes/launch |",
        "|---|---:|---:|---:|---:|",
        f"| FETCH_SIZE | {f['sweep_kib']:.1f} | {f['copy_kib']:.1f} | {f['factor']:.4f} | {f['bytes']:.0f} |",
        f"| WRITE_SIZE | {w['sweep_kib']:.1f} | {w['copy_kib']:.1f} | {w['factor']:.4f} | {w['bytes']:.0f} |",
        "",
        f"traffic_bytes_per_launch: {total:.0f}",
        "",
    ]
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines))
    print("\n".join(lines))


def _trace_rows(path):
    """(name, start ns, end ns, grid x) of every kernel dispatch of a
    rocprofv3 --kernel-trace output: the CSV kernel trace or the rocpd
    SQLite database."""
    if path.endswith(".csv"):
        import csv
        with open(path) as fh:
            return [(r["Kernel_Name"], int(r["Start_Timestamp"]), int(r["End_Timestamp"]), int(r["Grid_Size_X"]))
                    for r in csv.DictReader(fh)]
    import sqlite3
    con = sqlite3.connect(path)
    rows = con.execute("select name, start, end, grid_x from kernels").fetchall()
    con.close()
    return rows


def trace_summary(db, title, command, top=30):
    """Markdown summary of a rocprofv3 --kernel-trace run: per-kernel totals;
    the finest-level sweep's mean over full launches (launches queued after a
    solve has converged exit at once and are left out) beside its mean over
    every launch; and, for the refinement mode, the split of the median
    Krylov iteration (between consecutive k_ir_update_p launches)."""
    import statistics
    rows = sorted(_trace_rows(db), key=lambda r: r[1])
    tot = sum(e - s for _, s, e, _ in rows) / 1e6
    agg = {}
    for name, s0, e0, g in rows:
        a = agg.setdefault(name, [0, 0.0, float("inf"), 0.0, 0])
        d = (e0 - s0) / 1e3
        a[0] += 1
        a[1] += d
        a[2] = min(a[2], d)
        a[3] = max(a[3], d)
        a[4] = max(a[4], g)
    lines = [f"# {title}", "", "Command (one MI355X; `cd /tmp && export TMPDIR=/tmp` first; `<repo>` the repository "
             "root, `<out>` an output folder):", "",
             "```", command, "```", "", f"Total kernel time: {tot:.2f} ms", "",
             "| kernel | calls | total ms | avg us | min us | max us | max grid |", "|---|---:|---:|---:|---:|---:|---:|"]
    for name, (n, t, lo, hi, g) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        short = name.replace("tpamr::", "").replace("(anonymous namespace)::", "").split("(")[0]
        if len(short) > 90:
            short = short[:87] + "..."
        lines.append(f"| `{short}` | {n} | {t / 1e3:.2f} | {t / n:.2f} | {lo:.2f} | {hi:.2f} | {g} |")
    # the finest-level plain sweep (the template's trailing zero-start flag off)
    target = SWEEP_KERNEL.replace(" ", "")[:-1] + ",false>"
    sweep = [(e0 - s0) / 1e3 for name, s0, e0, g in rows
             if name.replace(" ", "").replace("void", "", 1).replace("tpamr::", "").split("(")[0] == target]
    if sweep:
        med = statistics.median(sweep)
        full = [d for d in sweep if d >= 0.5 * med]
        lines += ["", f"`{SWEEP_KERNEL}` (finest level): mean {statistics.mean(sweep):.2f} us over all "
                      f"{len(sweep)} launches, {statistics.mean(full):.2f} us over the {len(full)} full launches "
                      f"(at least half the median), median {med:.2f} us."]
    # the face and cell passes of a projection step around its solve
    passes = ("k_mixture", "k_gravity_potential", "k_gravity_jump", "k_divergence", "k_projection_rhs",
              "k_face_jump", "k_metric_apply", "k_face_scale", "k_axpby")
    found = [(p, agg[n]) for p in passes for n in agg
             if n.replace("tpamr::", "").split("(")[0].strip() == p]
    if found:
        lines += ["", "Face and cell passes around the solve (coefficients, jumps, force, divergence, correction):",
                  "", "| kernel | calls | avg us |", "|---|---:|---:|"]
        lines += [f"| `{p}` | {a[0]} | {a[1] / a[0]:.2f} |" for p, a in found]
    ends = [i for i, r in enumerate(rows) if "k_ir_update_p" in r[0]]
    its = []
    for i0, i1 in zip(ends, ends[1:]):
        seg = rows[i0 + 1:i1 + 1]
        wall = seg[-1][2] - rows[i0][2]
        if wall < 5e6:
            its.append((wall, seg))
    if its:
        wall, seg = sorted(its, key=lambda t: t[0])[len(its) // 2]
        busy = sum(e0 - s0 for _, s0, e0, _ in seg)
        lines += ["", f"Median refinement-mode Krylov iteration ({len(its)} iterations): {wall / 1e3:.1f} us from "
                      f"the end of one k_ir_update_p to the next, kernels busy {busy / 1e3:.1f} us, "
                      f"{len(seg)} launches, in order:", "", "| us | kernel |", "|---:|---|"]
        for name, s0, e0, _ in seg:
            short = name.replace("tpamr::", "").replace("(anonymous namespace)::", "").split("(")[0]
            lines.append(f"| {(e0 - s0) / 1e3:.1f} | `{short[:80]}` |")
    print("\n".join(lines))


def uniform_sweep(n, reps=50):
    """Finest-level Jacobi sweep on a uniform n x n mesh (16x16 blocks, walls:
    no irregular rows) in FP32 and FP64, beside the library's streaming
    passes (a copy, and five reads per write): the streaming rate the regular
    stencil path reaches. One JSON line."""
    import torch
    import tpamr
    torch.cuda.init()  # torch's HIP runtime first (it finds no GPU once the library has opened it)
    mesh = tpamr.BlockTreeMesh(base_mesh=(n, n), block_size=(16, 16), periodic=(False, False))
    out = {"workload": f"uniform{n}x{n}_block16", "cells": mesh.ncells}
    for prec in ("fp32_refine", "fp64"):
        s = tpamr.EllipticSolver(tpamr.PoissonOperator(mesh, np.ones(mesh.nfaces)), precision=prec)
        ms, nbytes = s.time_smoother(reps=reps)
        out[prec] = {"sweep_us": ms * 1e3, "compulsory_bytes": nbytes, "GBps": nbytes / (ms * 1e-3) / 1e9,
                     "irregular_nnz": s.level_info(0)["irregular_nnz"]}
        s.close()
    out["stream"] = stream_rates()
    print(json.dumps(out), flush=True)


def stream_rates(n=1 << 28, reps=10):
    """HBM streaming rates (bytes read plus written per second, GB/s) of the
    library's float4 pass over n-float arrays, timed with HIP events on the
    context stream: a plain copy (working set 2n floats) and the Jacobi
    sweep's mix of five reads per write (6n floats); n = 2^28 puts both far
    beyond the 256 MB Infinity Cache."""
    import ctypes as C
    from tpamr import _lib as L
    ctx = L.Context(0, L.TPAMR_FP32_REFINE)
    out = {}
    try:
        for name, reads in (("copy", 1), ("sweep_mix", 5)):
            ms, nb = C.c_double(), C.c_double()
            L.check(L.lib().tpamr_time_stream(ctx.h, n, reads, reps, C.byref(ms), C.byref(nb)), ctx.h)
            out[name] = dict(GBps=nb.value / (ms.value * 1e-3) / 1e9, bytes=nb.value, ms=ms.value)
    finally:
        ctx.close()
    return out


def committed_traffic():
    """Bytes per launch recorded by the last --pmc-summary, or None."""
    try:
        with open(TRAFFIC_MD) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        return None
    return None


def wave_main(args):
    """Stokes-wave-like run: [0, 2] x [0, 1], 32x32 blocks, levels 5-11 near
    the wave surface, density (1000, 1.2), gravity (0, -9.81); `args.steps`
    steps (default 10) of gravity! plus project! with the interface at each
    step's time (dt = 1e-4, the surface following its linear standing-wave
    phase), pressures checkpointed, J = kinetic energy after the last step
    and its gradient with respect to the first step's pressure rhs by the
    adjoint run. The run keeps every field in HBM (DeviceProjectionRun) and is
    timed in the headline mode and in all-FP64; `value` is the headline
    mode's solve throughput (cells x steps / summed device solve time)."""
    import torch
    import tpamr
    from tpamr import workloads as W
    t0 = time.perf_counter()
    dev = torch.device("cuda:0")
    mesh = W.wave_mesh(tuple(args.levels))
    mesh_s = time.perf_counter() - t0
    log(f"wave: {mesh.ncells} cells, {mesh.nblocks} blocks ({mesh_s:.1f} s)")
    setup_s = {}
    nsteps, dt = args.steps, 1e-4
    u0 = 0.01 * W.smooth_velocity(mesh, device="cuda:0")
    d_u0 = torch.as_tensor(u0, device=dev)
    d_ifaces = frac_end = None
    res, runs = {}, {}
    grad = None
    for prec in (HEADLINE, "fp64"):
        t1 = time.perf_counter()
        s = tpamr.EllipticSolver(tpamr.PoissonOperator(mesh, 1.0), precision=prec)
        setup_s[prec] = time.perf_counter() - t1
        log(f"{prec}: hierarchy {[lv['ncells'] for lv in s.levels]} ({setup_s[prec]:.1f} s)")
        if d_ifaces is None:
            # the interface of every step (fraction, apertures, centroid
            # offsets) in HBM: each timed step computes its coefficients and
            # gravity jumps on the GPU
            t1 = time.perf_counter()
            d_ifaces = []
            for n, data in enumerate(W.wave_interfaces(mesh, nsteps, dt)):
                if n == nsteps:
                    frac_end = data[0]
                    break
                d_ifaces.append(tuple(torch.as_tensor(a, device=dev) for a in data))
            log(f"interface data of {nsteps} steps ({time.perf_counter() - t1:.1f} s)")
        w = torch.as_tensor(tpamr.stepping.kinetic_energy_weights(mesh, frac_end, W.WAVE_DENSITY), device=dev)
        run = tpamr.DeviceProjectionRun(s, dt, rtol=args.rtol, adjoint_rtol=args.rtol)
        phys = dict(density=W.WAVE_DENSITY, gravity=W.WAVE_GRAVITY)
        run.forward_interface(d_u0, d_ifaces[:1], **phys)  # first device refresh uploads the refresh plan
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        uN = run.forward_interface(d_u0, d_ifaces, **phys)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        solve_s = sum(st.solve_ms for st in run.stats) * 1e-3
        J = float(0.5 * torch.sum(w * uN * uN))
        r = dict(mcells=mesh.ncells * nsteps / solve_s / 1e6, solve_ms_per_step=solve_s / nsteps * 1e3,
                 wall_s=wall, wall_ms_per_step=wall / nsteps * 1e3, cycles=[st.cycles for st in run.stats],
                 converged=all(st.converged for st in run.stats), J=J)
        log(f"{prec}: forward {r}")
        if prec == HEADLINE:
            t1 = time.perf_counter()
            g = run.backward(w * uN)
            torch.cuda.synchronize()
            r["adjoint_wall_s"] = time.perf_counter() - t1
            ast = g["stats"]
            r["adjoint_mcells"] = mesh.ncells * nsteps / (sum(st.solve_ms for st in ast) * 1e-3) / 1e6
            r["adjoint_cycles"] = [st.cycles for st in ast]
            r["adjoint_converged"] = all(st.converged for st in ast)
            grad = g["rhs"].cpu().numpy()
            r["grad_rhs_l2"] = float(np.linalg.norm(grad))
            log(f"{prec}: adjoint {r['adjoint_wall_s']:.1f} s, {r['adjoint_mcells']:.1f} Mcells/s")
            runs[prec] = [c.cpu().numpy() for c in run.checkpoints]
            if args.wave_fd:
                # J is quadratic in the rhs: central differences are exact for any step
                steps = [(k.cpu().numpy(), j.cpu().numpy()) for k, j in run.steps]
                host = tpamr.ProjectionRun(s, dt, rtol=1e-12, adjoint_rtol=args.rtol)
                d = W.smooth_velocity(mesh, seed=W.SEED + 3, device="cuda:0")
                d = tpamr.divergence(s, d)
                eps = float(np.max(np.abs(tpamr.divergence(s, u0)))) / dt / float(np.max(np.abs(d)))
                jp = tpamr.kinetic_energy(mesh, host.forward(u0, steps, rhs_shift=eps * d), frac_end,
                                          W.WAVE_DENSITY)[0]
                jm = tpamr.kinetic_energy(mesh, host.forward(u0, steps, rhs_shift=-eps * d), frac_end,
                                          W.WAVE_DENSITY)[0]
                fd = (jp - jm) / (2 * eps)
                r["fd_check"] = dict(fd=fd, adjoint=float(grad @ d), rel=abs(fd - grad @ d) / abs(fd))
                log(f"{prec}: finite differences {r['fd_check']}")
        else:
            runs[prec] = [c.cpu().numpy() for c in run.checkpoints]
        res[prec] = r
        del run
        s.close()
        torch.cuda.empty_cache()
    perr = max(float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for a, b in zip(runs[HEADLINE], runs["fp64"]))
    head, f64 = res[HEADLINE], res["fp64"]
    out = {
        "metric": METRIC, "value": head["mcells"], "unit": "Mcells/s", "n_gpus": 1, "steps": nsteps,
        "warmup": 0, "ms_per_step": head["solve_ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 V-cycle and inner CG vectors / f64 refinement, inner products",
        "data": "synthetic",
        "config": {"workload": f"stokes_wave_levels{args.levels[0]}-{args.levels[1]}_rho1000-1.2",
                   "cells": mesh.ncells, "blocks": mesh.nblocks, "block_size": list(mesh.block_size),
                   "dt": dt, "rtol": args.rtol, "precision": HEADLINE},
        "fp64_solver": f64, "headline": head,
        "speedup_vs_fp64": head["mcells"] / f64["mcells"],
        "host_setup_s": {"mesh": mesh_s, "solver": setup_s, "threads": min(16, os.cpu_count() or 1),
                         "note": "tree mesh and faces; then per mode the hierarchy build, first coefficient "
                                 "refresh and upload"},
        "checkpoint_pressure_error_vs_fp64": perr,
        "kinetic_energy_rel_diff_vs_fp64": abs(head["J"] - f64["J"]) / abs(f64["J"]),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def taylor_green_run(n=128, block=16, steps=2, warmup=1, nu=0.02, dt=1e-3, tol=1e-12, oracle_check=True):
    """The plain-method Taylor-Green configuration: periodic [0, 2 pi]^2, n^2
    cells in block^2 blocks, density 1, viscosity nu, dt: a fresh system
    runs `warmup` untimed one-velocity steps (step! with velocity and
    pressure in HBM) and then `steps` timed ones. With oracle_check, the
    oracle runs all of them on the host (its FP64 restatement: numpy
    transport and viscous operator, C++ pressure solver) and the velocity
    and pressure are compared field for field; its wall time per step is the
    CPU figure."""
    import torch
    import tpamr
    kw = dict(base_mesh=(n, n), block_size=(block, block), extent=(2 * np.pi, 2 * np.pi), periodic=(True, True))
    mesh = tpamr.BlockTreeMesh(**kw)
    O = _oracle()
    om = O.OMesh(**kw)
    ax = np.asarray(om.face_axis)
    c = om.face_center
    u0 = np.where(ax == 0, np.sin(c[:, 0]) * np.cos(c[:, 1]), -np.cos(c[:, 0]) * np.sin(c[:, 1]))
    system = tpamr.NavierStokesSystem(mesh, density=1.0, dynamic_viscosity=nu)

    def run(k, untimed=0):
        du = torch.tensor(u0, device="cuda:0")
        dp = torch.zeros(mesh.ncells, dtype=torch.float64, device="cuda:0")
        for _ in range(untimed):
            tpamr.step_device(du, dp, system, dt, tolerance=tol)
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = None
        for _ in range(k):
            st = tpamr.step_device(du, dp, system, dt, tolerance=tol)
        torch.cuda.synchronize()
        return time.perf_counter() - t, du, dp, st

    run(steps)
    # a new system starts the transport's sweep order forward, as the
    # oracle's does; its first `warmup` steps (context set-up, first graph
    # captures) are not timed, and the oracle runs them too
    system.close()
    system = tpamr.NavierStokesSystem(mesh, density=1.0, dynamic_viscosity=nu)
    wall, du, dp, st = run(steps, untimed=warmup)
    vx, vy = system.viscous_stats()
    out = {"cells": mesh.ncells, "block_size": [block, block], "steps": steps, "untimed_first_steps": warmup,
           "dt": dt, "viscosity": nu,
           "ms_per_step": wall / steps * 1e3, "mcells": mesh.ncells * steps / wall / 1e6,
           "pressure_cycles": st.cycles, "viscous_cycles": [vx.cycles, vy.cycles], "converged": bool(st.converged)}
    system.close()
    if oracle_check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_ns as N
        model, threads = host_cpu()
        O.set_threads(threads)
        S = O.OSolver(om, np.ones(om.nfaces))
        adv, D = N.StaggeredAdvection(om), N.StaggeredDiffusion(om, nu)
        uo, po = u0.copy(), np.zeros(om.ncells)
        t = time.perf_counter()
        for _ in range(warmup + steps):
            uo, po, _ = N.step(uo, po, S, adv, D, dt, tolerance=tol)
        cpu = (time.perf_counter() - t) / (warmup + steps)
        ug, pg = du.cpu().numpy(), dp.cpu().numpy()
        out["velocity_error_vs_oracle"] = float(np.max(np.abs(ug - uo)) / np.max(np.abs(uo)))
        out["pressure_error_vs_oracle"] = float(np.max(np.abs(pg - po)) / np.max(np.abs(po)))
        out["cpu_oracle_ms_per_step"] = cpu * 1e3
        out["cpu_oracle_threads"] = threads
        out["cpu_oracle_note"] = ("oracle restatement (numpy transport and viscous operator with an exact sparse "
                                  "viscous solve, C++ pressure solver on the host's threads), FP64")
    return out


def taylor_green_main(args):
    n = args.tg_cells
    res = taylor_green_run(n=n, steps=args.steps, warmup=args.warmup, oracle_check=n <= 512)
    out = {"metric": "Mcells/s of the one-velocity step (cells x steps / wall time)", "value": res["mcells"],
           "unit": "Mcells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"taylor_green_{n}x{n}_periodic", "cells": res["cells"], "viscosity": 0.02,
                      "dt": 1e-3},
           "detail": res}
    print(json.dumps(out), flush=True)


def cmdgfm_run(n=2048, steps=5, warmup=1, tol=1e-8, precision="fp64", sample=256):
    """The CMDGFM reduced-pressure/jump block solve on the synthetic
    n^2-cell system of tpamr.workloads.cmdgfm_block_workload (walls, liquid
    disk r=0.27 at the centre, 1000:1, dt 0.02): `steps` timed solves from
    the rhs and solution in HBM (the device entry point) after `warmup`. The
    oracle solves the same workload at sample^2 cells on the host and the
    GPU's answer there is compared with it (iterations, relative error); the
    oracle's solve rate at that size is the CPU figure."""
    import torch
    import tpamr
    from tpamr.workloads import cmdgfm_block_workload

    def system(mesh, kappa, sysd, prec):
        solver = tpamr.EllipticSolver(tpamr.PoissonOperator(mesh, kappa), precision=prec)
        gsys = tpamr.CMDGFMBlockSystem(solver, sysd["matrix"], sysd["pressure_indices"], sysd["pressure_anchors"],
                                       sysd["pressure_volumes"], sysd["jump_lengths"], omitted=sysd["omitted"],
                                       omitted_rhs=sysd["omitted_rhs"], omitted_volumes=sysd["omitted_volumes"])
        return solver, gsys

    def build(k):
        t = time.perf_counter()
        mesh, kappa, sysd, rhs, dt, rho = cmdgfm_block_workload(n=k)
        solver, gsys = system(mesh, kappa, sysd, precision)
        log(f"cmdgfm {k}^2: {mesh.ncells} cells, {len(sysd['jump_lengths'])} jumps, setup "
            f"{time.perf_counter() - t:.1f} s")
        return mesh, kappa, sysd, rhs, dt, rho, solver, gsys

    def timed(gsys, d_rhs, d_x):
        st = None
        for _ in range(warmup):
            st = gsys.solve_device(d_rhs, d_x, dt=dt, density=rho, tolerance=tol)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            st = gsys.solve_device(d_rhs, d_x, dt=dt, density=rho, tolerance=tol)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps, st

    mesh, kappa, sysd, rhs, dt, rho, solver, gsys = build(n)
    d_rhs = torch.tensor(rhs, device="cuda:0")
    d_x = torch.zeros(len(rhs), dtype=torch.float64, device="cuda:0")
    wall, st = timed(gsys, d_rhs, d_x)
    out = {"cells": mesh.ncells, "unknowns": len(rhs), "jumps": st.jump_unknowns, "ms_per_solve": wall * 1e3,
           "mcells": mesh.ncells / wall / 1e6, "iterations": st.iterations, "krylov_runs": st.krylov_runs,
           "converged": st.converged, "pressure_defect": st.pressure_defect, "jump_defect": st.jump_defect,
           "anchor_defect": st.anchor_defect, "precision": precision}
    solver.close()
    if precision == "fp64":
        # the pressure block's V-cycle in FP32 (BiCGSTAB, products and defects stay FP64)
        x64 = d_x.clone()
        solver, gsys = system(mesh, kappa, sysd, "fp32_vcycle")
        wall, st = timed(gsys, d_rhs, d_x)
        out["fp32_vcycle"] = {"ms_per_solve": wall * 1e3, "mcells": mesh.ncells / wall / 1e6,
                              "iterations": st.iterations, "converged": st.converged,
                              "pressure_defect": st.pressure_defect,
                              "solution_vs_fp64": float((d_x - x64).abs().max() / x64.abs().max())}
        solver.close()
    if sample:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_cmdgfm as OC
        O = _oracle()
        model, threads = host_cpu()
        O.set_threads(threads)
        mesh, kappa, sysd, rhs, dt, rho, solver, gsys = build(sample)
        x, gst = gsys.solve(rhs, dt=dt, density=rho, tolerance=tol)
        solver.close()
        om = O.OMesh(base_mesh=(sample, sample), block_size=(16, 16), periodic=(False, False))
        osol = O.OSolver(om, kappa)
        osys = OC.BlockSystem(sysd["matrix"], len(sysd["pressure_indices"]), sysd["pressure_indices"],
                              sysd["pressure_anchors"], sysd["pressure_volumes"], sysd["jump_lengths"],
                              sysd["omitted"], sysd["omitted_rhs"], sysd["omitted_volumes"], mesh.ncells,
                              osol.vcycle)
        t = time.perf_counter()
        ox, ost = OC.solve(osys, rhs, dt=dt, density=rho, reference_length=1.0, tolerance=tol)
        cpu = time.perf_counter() - t
        out["sample"] = {"cells": mesh.ncells, "gpu_iterations": gst.iterations, "oracle_iterations": ost["iterations"],
                         "error_vs_oracle": float(np.max(np.abs(x - ox)) / np.max(np.abs(ox))),
                         "cpu_oracle_ms": cpu * 1e3, "cpu_oracle_mcells": mesh.ncells / cpu / 1e6,
                         "cpu_threads": threads, "cpu_model": model,
                         "cpu_note": "oracle restatement: numpy/SciPy BiCGSTAB, block preconditioner and Crout ILU, "
                                     "C++ V-cycle on the host's threads, FP64"}
    return out


def cmdgfm_main(args):
    n = args.cm_cells
    res = cmdgfm_run(n=n, steps=args.steps, warmup=args.warmup, sample=args.cm_sample)
    sample = res.get("sample", {})
    out = {"metric": "Mcells/s of the CMDGFM block solve (pressure cells / solve wall time)", "value": res["mcells"],
           "unit": "Mcells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": res["ms_per_solve"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"cmdgfm_block_{n}x{n}_disk_rho1000", "cells": res["cells"],
                      "jumps": res["jumps"], "tolerance": 1e-8},
           "cpu_baseline": None if not sample else {
               "value": sample["cpu_oracle_mcells"], "unit": "Mcells/s", "cores": sample["cpu_threads"],
               "kind": "port", "sample": f"oracle block solve of the same workload at {sample['cells']} cells"},
           "detail": res}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--levels", type=int, nargs=2, default=(5, 10))
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-levels", type=int, nargs=2, default=(5, 8))
    ap.add_argument("--cpu-full", action="store_true",
                    help="also solve the benchmark system itself with the oracle on the host's threads and "
                         "report every mode's pressure error against it (minutes of CPU setup)")
    ap.add_argument("--probe", action="store_true", help="short run for rocprofv3 --pmc passes")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DB", "WRITE_DB"))
    ap.add_argument("--trace-summary", nargs=3, metavar=("DB", "TITLE", "COMMAND"))
    ap.add_argument("--workload", choices=("circles", "wave", "taylor_green", "cmdgfm"), default="circles",
                    help="circles: the headline solve benchmark; wave: the checkpointed multi-step run; "
                         "taylor_green: one-velocity steps on the periodic Taylor-Green field; "
                         "cmdgfm: the CMDGFM reduced-pressure/jump block solve")
    ap.add_argument("--tg-cells", type=int, default=128, help="taylor_green: cells per side")
    ap.add_argument("--cm-cells", type=int, default=2048, help="cmdgfm: cells per side")
    ap.add_argument("--cm-sample", type=int, default=256,
                    help="cmdgfm: cells per side of the oracle comparison and CPU figure (0: none)")
    ap.add_argument("--wave-fd", action="store_true", help="wave: one finite-difference check of the gradient")
    ap.add_argument("--uniform", type=int, metavar="N",
                    help="time the finest-level sweep on a uniform N x N mesh of 16x16 blocks (no irregular rows)")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.trace_summary:
        trace_summary(*args.trace_summary)
        return
    if args.uniform:
        uniform_sweep(args.uniform)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    if args.probe:
        probe(args.levels)
        return
    if args.workload == "taylor_green":
        taylor_green_main(args)
        return
    if args.workload == "cmdgfm":
        cmdgfm_main(args)
        return
    if args.workload == "wave":
        if "--levels" not in sys.argv:
            args.levels = (5, 11)
        if "--steps" not in sys.argv:
            args.steps = 10
        wave_main(args)
        return
    import torch

    t0 = time.perf_counter()
    results = {}
    x_sol = {}
    case = Case(args.levels)
    mesh = case.mesh
    workload_s = time.perf_counter() - t0
    log(f"workload: {mesh.ncells} cells ({workload_s:.1f} s)")
    setup_s = {}
    for prec in MODES:
        t1 = time.perf_counter()
        solver = case.solver(prec)
        d_rhs = case.d_rhs
        setup_s[prec] = time.perf_counter() - t1
        log(f"{prec}: hierarchy {[lv['ncells'] for lv in solver.levels]} ({setup_s[prec]:.1f} s)")
        d_x = torch.zeros(mesh.ncells, dtype=torch.float64, device="cuda:0")
        wall, dev, st = timed_solves(solver, d_rhs, d_x, args.steps, args.warmup, args.rtol)
        x_sol[prec] = d_x.cpu().numpy()
        res = dict(ms_per_step=wall * 1e3, device_ms=dev * 1e3, mcells=mesh.ncells / wall / 1e6,
                   cycles=st.cycles, converged=st.converged, final_linf=st.final_linf,
                   target_linf=st.target_linf, initial_linf=st.initial_linf)
        if prec == HEADLINE:
            ms, nbytes = solver.time_smoother(reps=50)
            info = solver.level_info(0)
            res["smoother_ms"] = ms
            res["smoother_kernel_bytes"] = nbytes
            res["smoother_bytes"] = sweep_algorithmic_bytes(mesh.ncells, info["irregular_nnz"])
            res["irregular_nnz"] = info["irregular_nnz"]
            res["levels"] = len(solver.levels)
        results[prec] = res
        log(f"{prec}: {res}")
        solver.close()
        del d_x
        torch.cuda.empty_cache()
    log(f"timed workload done ({time.perf_counter() - t0:.1f} s)")

    head, f32v, f64 = results[HEADLINE], results["fp32_vcycle"], results["fp64"]

    def perr(prec):
        return float(np.max(np.abs(x_sol[prec] - x_sol["fp64"])) / np.max(np.abs(x_sol["fp64"])))

    achieved = head["smoother_bytes"] / (head["smoother_ms"] * 1e-3) / 1e9
    stream = stream_rates()
    traffic = committed_traffic()
    out = {
        "metric": METRIC,
        "value": head["mcells"],
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle and inner CG vectors / f64 refinement, inner products",
        "data": "synthetic",
        "config": {"workload": f"circles64_levels{args.levels[0]}-{args.levels[1]}_rho1000",
                   "cells": mesh.ncells, "blocks": mesh.nblocks, "block_size": list(mesh.block_size),
                   "mg_levels": head["levels"], "rtol": args.rtol, "density_ratio": 1000.0,
                   "precision": HEADLINE},
        "cycles": head["cycles"],
        "converged": bool(head["converged"]),
        "fp64_solver": {"value": f64["mcells"], "ms_per_step": f64["ms_per_step"], "cycles": f64["cycles"],
                        "converged": bool(f64["converged"])},
        "fp32_vcycle_solver": {"value": f32v["mcells"], "ms_per_step": f32v["ms_per_step"],
                               "cycles": f32v["cycles"], "converged": bool(f32v["converged"]),
                               "pressure_error_vs_fp64_gpu": perr("fp32_vcycle")},
        "speedup_vs_fp64": head["mcells"] / f64["mcells"],
        "host_setup_s": {"workload": workload_s, "solver": setup_s,
                         "threads": min(16, os.cpu_count() or 1),
                         "note": "mesh, fraction and velocity; then per mode the hierarchy build, first "
                                 "coefficient refresh and upload (the first mode also computes the rhs)"},
        "pressure_error_vs_fp64_gpu": perr(HEADLINE),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": SWEEP_KERNEL, "bytes_per_launch": head["smoother_bytes"],
                     "kernel_compulsory_bytes": head["smoother_kernel_bytes"],
                     "launch_ms": head["smoother_ms"], "copy_GBps": stream["copy"]["GBps"],
                     "stream_GBps": stream["sweep_mix"]["GBps"],
                     "frac_of_stream": achieved / stream["sweep_mix"]["GBps"],
                     "stream_note": "hand-written float4 passes over 2^28-float arrays (beyond the Infinity "
                                    "Cache): copy, and five reads per write as the sweep"},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"], out["cpu_sample_on_gpu"] = cpu_baseline(tuple(args.cpu_levels), args.rtol)
    else:
        out["cpu_baseline"] = None
    if args.cpu_full:
        out["cpu_full"] = cpu_full(case, x_sol, args.rtol)
    # the plain-method configuration's one-velocity step (128^2 cells, two
    # steps), field for field against the oracle
    out["taylor_green_step"] = taylor_green_run(steps=4, warmup=1)
    # the CMDGFM reduced-pressure/jump block solve (2048^2 cells), with the
    # oracle's answer and rate at 256^2
    out["cmdgfm_block_solve"] = cmdgfm_run(n=2048, steps=3, warmup=1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
