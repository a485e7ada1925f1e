#!/usr/bin/env python3
"""Benchmark: the SIMPLEC pressure-correction solve on one MI355X.

Workload (one step = one pressure solve to 1e-6 relative residual): the
pressure-correction operator of a 55x55x55 Freudenthal tetrahedral box on
[0,1]^3 (998,250 cells, 4,954,950 nonzeros), interior points perturbed by
seeded uniform noise of +-0.2 h, renumbered by reverse Cuthill-McKee, FP64
geometry, face coefficients |S|^2/(S.d) assembled on the device by
fv_pressure_assemble with the cell at the origin pinned (the SIMPLEC
closed-domain default), and a seeded zero-mean Gaussian right-hand side.
Aggregation AMG (pairwise, 2 passes, symmetric M-matrix coarsening) as the
preconditioner of device-resident PCG, x0 = 0, FP32 arithmetic with FP64
reductions. Setup (mesh, renumbering, aggregation) is outside the timed
region; operator and right-hand side are resident in HBM when timing starts.

Prints one JSON line with the metric, the roofline figure of the fine-level
l1-Jacobi smoothing kernel (the largest cost) and the serial CPU oracle timed
on one host core beside it.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--sweeps 5] [--tail T] [--no-cpu-baseline]
    python bench.py --copy-probe BYTES [--steps K]   # HBM counter calibration
    python bench.py --fv-kernels [--steps K]         # each FV kernel at C2 size
    python bench.py --simplec-box N [--steps K]       # SIMPLEC on the N^3 lid-driven box

The line also carries "large_mesh": the plain path and the K-cycle variant
timed on the 110^3 box (7,986,000 cells), and "simplec": ms per steady SIMPLEC iteration on the
128x128x1 lid-driven cavity at Re=1000 (500 iterations, iterations 6-500
timed), with the oracle timed on one core beside it, and "simplec_sst": ms
per SST k-omega SIMPLEC iteration on the 4.06M-cell backward-facing step
(iterations 4-13 timed), with the oracle timed on one core on the same step
with 2 spanwise cells.

    python bench.py --simplec-step NZ [--steps K]    # SST step with NZ spanwise cells only
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "ms/SIMPLEC iteration & AMG-PCG ms to 1e-6 rel res, learned vs plain; HBM GB/s"
HBM_PEAK_GBS = 8000.0
SEED = 20261202


def rhs_for(n: int, pinned: int, seed: int = SEED):
    """Zero-mean Gaussian RHS: numpy PCG64(seed) normals, mean removed in FP64,
    rounded to FP32, zero at the pinned cell."""
    import numpy as np
    b = np.random.default_rng(seed).standard_normal(n)
    b -= b.mean()
    b = b.astype(np.float32)
    if pinned >= 0:
        b[pinned] = 0
    return b


def fp64_reference(problem, b, x0, solve_fp32):
    """The FP64 solution of the same FP32-valued operator, by iterative
    refinement: residuals of b - A x in FP64 (scipy, CPU), corrections from
    the GPU's FP32 AMG-PCG on the scaled residual, accumulated in FP64, until
    the FP64 relative residual is below 1e-13. Returns (x64, its relative
    residual, refinement steps)."""
    import numpy as np
    import scipy.sparse as sp
    a = sp.csr_matrix((problem["values"].astype(np.float64), problem["col"], problem["row_ptr"]))
    b64 = b.astype(np.float64)
    bn = np.linalg.norm(b64)
    x = x0.astype(np.float64)
    steps = 0
    rel = np.linalg.norm(b64 - a @ x) / bn
    while rel > 1e-13 and steps < 8:
        r = b64 - a @ x
        scale = np.max(np.abs(r))
        x += solve_fp32((r / scale).astype(np.float32)).astype(np.float64) * scale
        steps += 1
        rel = np.linalg.norm(b64 - a @ x) / bn
    return x, float(rel), steps


def fp64_error(x, x64) -> dict:
    import numpy as np
    d = x.astype(np.float64) - x64
    return {"l2": float(np.linalg.norm(d) / np.linalg.norm(x64)),
            "max_over_max": float(np.max(np.abs(d)) / np.max(np.abs(x64)))}


# The fast pressure variant (no reference counterpart as a configuration; the
# cycle is the reference's AmgCycle::K): flexible AMG-PCG whose coarse levels
# of at least 50,000 rows take the K-cycle's two inner Krylov corrections,
# plain V-cycle recursion below, one pre- and one post-smoothing sweep, a
# five-level hierarchy whose coarsest level gets 4 sweeps. Picked by timing the
# 1M- and 8M-cell solves over sweeps 1-5, K depths 1-4, 2-32 coarsest sweeps,
# 5-8 levels and row bounds of 20k-120k: the levels past the fifth barely
# coarsen and are latency-bound, and a level below ~50k rows costs more on a
# second visit than it saves.
KCYCLE = dict(pre_sweeps=1, post_sweeps=1, coarse_sweeps=4, max_levels=5, k_min_rows=50000)


def time_solver(dev, problem, options, kind, b, x, control, steps, warmup, graphs, direct=0, bf16=False,
                tail=-1):
    """Builds a solver on the problem's operator, refreshes it (timed), lets
    the autotuner plan its V-cycle (unless a dense level replaces the tail),
    and times `steps` solves of b from x = 0 after `warmup`. Returns the
    solver, its last stats and the figures."""
    import numpy as np

    import kutta_amd
    solver = kutta_amd.make_amg_solver(kind, dev, problem.host, options)
    level = solver.set_direct(direct) if direct else 0
    solver.set_coarse_bf16(bf16)
    solver.set_graphs(graphs)
    solver.set_tail(tail)
    t = time.perf_counter()
    solver.refresh(problem)
    dev.synchronize()
    refresh_ms = (time.perf_counter() - t) * 1e3
    if tail < 0 and not level:
        solver.autotune_tail(problem, 5)
    for _ in range(max(warmup, 1)):
        x.fill(np.float32(0))
        solver.solve(problem, x, b, control)
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        x.fill(np.float32(0))
        st = solver.solve(problem, x, b, control)
    dev.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return solver, st, {"ms": round(ms, 4), "iterations": st.iterations,
                        "relative_residual": st.final_residual / st.initial_residual, "converged": st.converged,
                        "refresh_ms": round(refresh_ms, 2), "refresh_plus_solve_ms": round(ms + refresh_ms, 4),
                        "direct_level": level, "hierarchy": solver.describe(), "plan": solver.plan()}


def lazy_dense_sequence(dev, problem, options, b, x, control, direct, graphs, steps=12, drift=1e-3,
                        tolerance=1e-2) -> dict:
    """The mixed path over a drifting operator, as SIMPLEC refreshes it every
    outer iteration: the operator's values scaled by 1 + drift*k at step k
    (untimed upload), then refresh + solve timed, for the plain FP32 path, the
    mixed path with the dense inverse recomputed at every refresh, and the
    mixed path with the lazy refresh (recomputed once the level moved more
    than `tolerance` from the values of its inverse). A uniform scaling is a
    simple model of the slow drift of a converging SIMPLEC run."""
    import numpy as np

    import kutta_amd
    from kutta_amd import SolverKind
    m = kutta_amd.upload(dev, problem.host)
    vals = m.val.download()
    out = {"model": f"values x (1 + {drift:g} k), k = 1..{steps}; refresh + solve per step",
           "tolerance": tolerance}
    for label, dr, tol in (("plain", 0, 0.0), ("mixed_eager", direct, 0.0), ("mixed_lazy", direct, tolerance)):
        m.val.upload(vals)
        solver = kutta_amd.make_amg_solver(SolverKind.AmgPcg, dev, problem.host, options)
        if dr:
            solver.set_direct(dr)
            solver.set_dense_tolerance(tol)
        solver.set_graphs(graphs)
        solver.refresh(m)
        if not dr:
            solver.autotune_tail(m, 5)
        x.fill(np.float32(0))
        solver.solve(m, x, b, control)
        total, iterations = 0.0, []
        for k in range(1, steps + 1):
            m.val.upload((vals * np.float32(1.0 + drift * k)).astype(np.float32))
            dev.synchronize()
            t0 = time.perf_counter()
            solver.refresh(m)
            x.fill(np.float32(0))
            st = solver.solve(m, x, b, control)
            dev.synchronize()
            total += time.perf_counter() - t0
            iterations.append(st.iterations)
        entry = {"refresh_plus_solve_ms": round(total * 1e3 / steps, 3), "iterations": iterations}
        if dr:
            inv, kept, _ = solver.dense_refreshes()
            entry.update({"dense_inversions": inv, "dense_kept": kept})
        out[label] = entry
        del solver
    out["lazy_vs_plain"] = round(out["plain"]["refresh_plus_solve_ms"] / out["mixed_lazy"]["refresh_plus_solve_ms"], 3)
    out["lazy_vs_eager"] = round(out["mixed_eager"]["refresh_plus_solve_ms"] /
                                 out["mixed_lazy"]["refresh_plus_solve_ms"], 3)
    m.val.upload(vals)
    return out


def large_mesh_comparison(dev, size, steps, warmup, graphs, args) -> dict:
    """The plain FP32 AMG-PCG path (V(5,5), the reference's configuration) and
    the K-cycle variant (KCYCLE) on the size^3 box (7,986,000 cells at 110): ms to
    1e-6, with and without the refresh each solve needs when the operator
    changes (every SIMPLEC outer iteration), and each solution's error
    against the FP64 solution of the same operator."""
    import numpy as np

    import kutta_amd
    from kutta_amd import AmgCycle, AmgOptions, OperatorStructure, PressureProblem, SolveControl, SolverKind
    t = time.perf_counter()
    problem = PressureProblem(dev, size, size, size, tet=True, amplitude=0.2, seed=SEED, reference=-2)
    setup_s = time.perf_counter() - t
    b_host = rhs_for(problem.n_cells, problem.reference)
    b, x = dev.upload(b_host), dev.alloc(problem.n_cells)
    control = SolveControl(rel_tol=1e-6, abs_tol=0.0, max_iter=2000)
    base = dict(aggregation_passes=2, structure=OperatorStructure.SymmetricMMatrix)
    plain, _, out_plain = time_solver(dev, problem, AmgOptions(pre_sweeps=5, post_sweeps=5, **base),
                                      SolverKind.AmgPcg, b, x, control, steps, warmup, graphs)
    x_plain = x.download()

    def gpu_solve(rhs):
        x.fill(np.float32(0))
        plain.solve(problem, x, dev.upload(rhs), control)
        return x.download()

    x64, x64_residual, _ = fp64_reference(problem, b_host, x_plain, gpu_solve)
    out_plain["fp64_error"] = fp64_error(x_plain, x64)
    del plain
    _, _, out_k = time_solver(dev, problem, AmgOptions(cycle=AmgCycle.K, **KCYCLE, **base), SolverKind.AmgPcg,
                              b, x, control, steps, warmup, graphs)
    out_k["fp64_error"] = fp64_error(x.download(), x64)
    out_k["speedup_vs_plain_pcg"] = round(out_plain["ms"] / out_k["ms"], 3)
    out_k["speedup_refresh_plus_solve"] = round(out_plain["refresh_plus_solve_ms"] / out_k["refresh_plus_solve_ms"], 3)
    print(f"[bench] {problem.n_cells} cells: plain {out_plain['ms']:.2f} ms ({out_plain['iterations']} it), "
          f"K-cycle {out_k['ms']:.2f} ms ({out_k['iterations']} it), {out_k['speedup_vs_plain_pcg']}x",
          file=sys.stderr, flush=True)
    return {"workload": f"the same pressure solve on the {size}^3 box ({problem.n_cells} cells, {problem.nnz} nnz)",
            "setup_s": round(setup_s, 2), "fp64_relative_residual": x64_residual,
            "plain_pcg": out_plain, "pcg_kcycle": out_k}


def learned_variant(dev, problem, options, b, x, x_plain, x64, control, plain_ms, direct, args) -> dict:
    """The learned coarse correction inside AMG-FGMRES(30) at the scale of its
    specification: the solver harvests (level residual, and after the solve
    its exact level correction by AMG-PCG) pairs from its own steps on a
    sequence of seeded right-hand sides until --learned-pairs are held; the
    network trains for --learned-epochs epochs on the first 90% and is scored
    on the held-out 10%; then b is solved with the stage under its gate
    (applied only if the last epoch left at most --learned-max-unexplained of
    the targets unexplained; dropped within a solve when the residual falls by
    less than 0.5 per step) and with the stage applied regardless."""
    import numpy as np

    import kutta_amd
    from kutta_amd import SolverKind
    n = problem.n_cells
    other = kutta_amd.make_amg_solver(SolverKind.AmgFgmres, dev, problem.host, options)
    level = other.set_direct(direct) if direct else 0
    other.refresh(problem)
    other.set_learned(level=args.learned_level, layers=3, seed=0, capacity=args.learned_pairs)
    other.set_learned_mode(1)
    harvest_s = targets_s = 0.0
    solves = pairs = 0
    while (solves < args.learned_harvest if args.learned_harvest is not None else
           other.learned_report()["harvested"] < args.learned_pairs and solves < 4 * args.learned_pairs):
        t = time.perf_counter()
        for _ in range(50):  # a chunk of right-hand sides, then their targets
            x.fill(np.float32(0))
            other.solve(problem, x, dev.upload(rhs_for(n, problem.reference, SEED + 1 + solves)), control)
            solves += 1
            if args.learned_harvest is not None and solves >= args.learned_harvest:
                break
        dev.synchronize()
        harvest_s += time.perf_counter() - t
        t = time.perf_counter()
        pairs = other.learned_targets(1e-6)
        dev.synchronize()
        targets_s += time.perf_counter() - t
        print(f"[bench] learned stage: {solves} harvest solves, {pairs} pairs with targets", file=sys.stderr,
              flush=True)
    train_pairs = max(1, (9 * pairs) // 10)
    steps = args.learned_train if args.learned_train else args.learned_epochs * train_pairs
    losses, unexplained = [], []
    t = time.perf_counter()
    while len(losses) < steps:  # one epoch per call; the Adam state carries over
        l, u = other.learned_train(min(train_pairs, steps - len(losses)), 1e-3, train_pairs)
        losses += list(l)
        unexplained += list(u)
        print(f"[bench] learned stage: {len(losses)} of {steps} Adam steps, unexplained "
              f"{float(np.mean(u)):.3f}", file=sys.stderr, flush=True)
    losses, unexplained = np.asarray(losses), np.asarray(unexplained)
    train_s = time.perf_counter() - t
    held_out = other.learned_evaluate(train_pairs, pairs - train_pairs) if pairs > train_pairs else None
    rep = other.learned_report()

    def timed(max_unexplained):
        other.set_learned_mode(2, 0.5, max_unexplained)
        for _ in range(max(args.warmup, 1)):
            x.fill(np.float32(0))
            other.solve(problem, x, b, control)
        dev.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            x.fill(np.float32(0))
            st = other.solve(problem, x, b, control)
        dev.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps, st, x.download()

    vms, vst, xv = timed(args.learned_max_unexplained)
    gate = other.learned_report()
    fms, fst, xf = timed(1e30)
    forced = other.learned_report()
    epochs = [float(np.mean(unexplained[e:e + train_pairs])) for e in range(0, len(unexplained), train_pairs)]
    return {"ms": round(vms, 4), "iterations": vst.iterations,
            "relative_residual": vst.final_residual / vst.initial_residual, "converged": vst.converged,
            "speedup_vs_plain_pcg": round(plain_ms / vms, 3), "direct_level": level,
            "learned_level": rep["level"], "learned_rows": rep["rows"], "pairs": pairs,
            "harvest_solves": solves, "train_pairs": train_pairs, "epochs": round(len(losses) / train_pairs, 3),
            "train_steps": len(losses), "unexplained_by_epoch": [round(e, 4) for e in epochs],
            "unexplained_held_out": held_out, "harvest_s": round(harvest_s, 2), "targets_s": round(targets_s, 2),
            "train_s": round(train_s, 2), "solves_in_learned_mode": gate["solves"], "gated": gate["gated"],
            "fallbacks": gate["fallbacks"], "fp64_error": fp64_error(xv, x64),
            "max_abs_diff_vs_plain_over_max_abs_x":
                float(np.abs(xv.astype(np.float64) - x_plain).max() / np.abs(x_plain).max()),
            "forced": {"ms": round(fms, 4), "iterations": fst.iterations, "converged": fst.converged,
                       "relative_residual": fst.final_residual / fst.initial_residual,
                       "applications": forced["applications"] - gate["applications"],
                       "fallbacks": forced["fallbacks"] - gate["fallbacks"], "fp64_error": fp64_error(xf, x64)}}


def cpu_baseline(problem, b, sweeps, gpu_stats):
    """The serial CPU oracle (the reference's CPU backend restated) solving the
    same operator bytes and RHS, pinned to one core."""
    import numpy as np

    import oracle
    a = oracle.Csr.from_arrays(problem["row_ptr"], problem["col"], problem["diag"],
                               problem["values"].astype(np.float64))
    solver = oracle.Solver("AmgPcg", a, pre=sweeps, post=sweeps, passes=2, m_matrix=True)
    previous = os.sched_getaffinity(0)
    core = min(previous)
    os.sched_setaffinity(0, {core})
    try:
        t = time.perf_counter()
        x, st = solver.solve(b, rel_tol=1e-6, abs_tol=0.0, max_iter=2000)
        seconds = time.perf_counter() - t
    finally:
        os.sched_setaffinity(0, previous)
    return {
        "value": round(seconds * 1e3, 3),
        "unit": "ms",
        "cores": 1,
        "kind": "port",
        "sample": (f"one full AMG-PCG V({sweeps},{sweeps}) solve of the same {a.n_rows}-row operator "
                   f"and RHS to 1e-6 ({st['iterations']} iterations), serial oracle pinned to one core"),
        "iterations": st["iterations"],
        "relative_residual": st["final_residual"] / st["initial_residual"],
        "iterations_match_gpu": st["iterations"] == gpu_stats.iterations,
    }, x


TRAFFIC_PROFILE = ROOT / "profiles" / "amg_smooth_fine_c2_hbm_traffic.md"


def measured_traffic():
    """HBM bytes per launch of the fine-level amg_smooth on the 55^3 operator,
    from the committed FETCH_SIZE/WRITE_SIZE counter summary (scaled by a copy
    kernel of known size), or None when that summary is absent."""
    try:
        for line in TRAFFIC_PROFILE.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_probe(nbytes: int, repeats: int) -> int:
    """Times `repeats` device-to-device copies of nbytes through the device ABI
    (kh_copy) and prints one JSON line. A copy moves exactly nbytes in and
    nbytes out of HBM, so profiler byte counters read on it calibrate the same
    counters read on the solver's kernels."""
    import numpy as np

    import kutta_amd
    dev = kutta_amd.Device(0)
    words = nbytes // 4
    src, dst = dev.alloc(words), dev.alloc(words)
    src.fill(np.float32(1))
    dev.copy(dst, src)
    dev.synchronize()
    t = time.perf_counter()
    for _ in range(repeats):
        dev.copy(dst, src)
    dev.synchronize()
    seconds = (time.perf_counter() - t) / repeats
    print(json.dumps({"probe": "device copy", "bytes": words * 4, "repeats": repeats,
                      "us_per_copy": round(seconds * 1e6, 3),
                      "GB/s (read+write)": round(2 * words * 4 / seconds / 1e9, 1)}))
    return 0


def fv_touched_bytes(mesh, domain: str, spec: str, args) -> int:
    """Compulsory bytes of one FV kernel launch: each array once, over the
    elements the launch can touch. Face arrays touch the launch's face range
    (the boundary slice for boundary-face kernels, the internal faces for
    internal-face kernels); a boundary-face kernel reaches only the cells
    that own boundary faces; everything else is touched whole."""
    import numpy as np
    owners = int(np.unique(mesh.owner[mesh.internal:]).size)
    face_like = {"F", "F+", "oF", "owner"}
    cell_like = {"C", "C+", "oC", "Cu", "mask", "oCi", "C6", "oC6"}
    total = 0
    for role, a in zip(spec.split(), args):
        if not isinstance(a, np.ndarray):
            continue
        per = a.nbytes // max(a.size, 1)
        width = 6 if role in ("C6", "oC6") else 1
        if role in face_like and domain == "boundary":
            count = mesh.boundary
        elif role in face_like and domain == "internal":
            count = mesh.internal
        elif role in cell_like and domain == "boundary":
            count = owners * width
        else:
            count = a.size
        total += count * per
    return total


def fv_kernel_bench(size: int, repeats: int) -> int:
    """Times every finite-volume kernel on the perturbed size^3 tetrahedral box
    (998,250 cells at size 55) with seeded inputs (kutta_amd.fv_specs) and
    prints one JSON line. Each figure is the wall time of `repeats`
    back-to-back launches on the device stream divided by `repeats`, so a
    kernel shorter than the launch interval reports the interval; the
    kernel-trace profile gives the execution time alone (--trace-summary with
    this line as --trace-line joins the two).
    Algorithmic bytes count each array argument once, over the elements the
    launch touches (fv_touched_bytes: a boundary-face kernel touches the
    boundary slice of face arrays and only the cells that own boundary
    faces); an array updated in place counts once. The roofline is HBM
    bandwidth."""
    import numpy as np

    import kutta_amd
    from kutta_amd.fv_specs import SPECS, FvMesh, build_args
    dev = kutta_amd.Device(0)
    mesh = FvMesh(kutta_amd.MeshSetup(size, size, size, tet=True, amplitude=0.2, seed=SEED))
    rows = []
    for name in sorted(SPECS):
        domain, spec = SPECS[name]
        n = mesh.domain(domain)
        args, outputs = build_args(mesh, spec, seed=3)
        on_device = [dev.upload(a) if isinstance(a, np.ndarray) else a for a in args]
        kernel = kutta_amd.Kernel[name]
        dev.launch(kernel, n, *on_device)
        dev.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            dev.launch(kernel, n, *on_device)
        dev.synchronize()
        seconds = time.perf_counter() - t0
        nbytes = fv_touched_bytes(mesh, domain, spec, args)
        us = seconds * 1e6 / repeats
        gbs = nbytes / (us * 1e-6) / 1e9
        rows.append({"kernel": name, "n": n, "bytes": nbytes, "us": round(us, 3),
                     "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
        print(f"[fv] {name:36s} n={n:8d} {us:9.3f} us {gbs:8.1f} GB/s", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "finite-volume kernel HBM bandwidth", "unit": "GB/s",
                      "peak": HBM_PEAK_GBS, "repeats": repeats, "data": "synthetic",
                      "config": {"workload": f"perturbed {size}^3 tet box, RCM ordered",
                                 "cells": mesh.cells, "faces": mesh.n_faces},
                      "kernels": rows}))
    return 0


def logical_traffic(dev, solver, ms: float, iterations: int = 3) -> dict:
    """The reference's logical-bytes model of one SIMPLEC iteration
    (bench/bench_fv.cpp:68-93: every buffer argument of every launch at its
    full size, plus 4n/8n/16n per reduction), counted by the runtime over
    `iterations` more iterations launched kernel by kernel, and that traffic
    at the timed ms per iteration as a fraction of the 8 TB/s HBM peak. The
    model counts whole buffers, so it overstates what a kernel touches; the
    reference reports it for the same comparison."""
    solver.set_graphs(False)
    dev.logical_bytes(True)
    solver.iterate(iterations)
    dev.synchronize()
    total, launches = dev.logical_bytes(False)
    per = total / iterations
    return {"model": "logical bytes (bench_fv.cpp:68-93)", "bytes_per_iteration": round(per),
            "bytes_per_cell": round(per / solver.info()["cells"], 1),
            "launches_per_iteration": round(launches / iterations, 1),
            "GB/s_at_timed_ms": round(per / (ms * 1e-3) / 1e9, 1),
            "frac": round(per / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def simplec_bench(dev, size: int, nz: int, reynolds: float, relaxation: float, monitor: int,
                  warmup: int, iterations: int, cpu_iterations: int, graphs: bool = True) -> dict:
    """ms per steady SIMPLEC outer iteration on the lid-driven cavity
    make_box_hex(size, size, nz): lid u=1 on ymax, walls at rest, symmetry on
    zmin/zmax, viscosity 1/Re, paired state and the reference's default
    numerics (validation/cases/cavity.hpp:111-121 for the 2-D cavity;
    bench/bench_fv.cpp:128-143 for the 3-D box). `warmup` iterations run
    untimed (setup, first monitors), then `iterations` are timed on the device
    stream. The oracle's serial restatement is timed on one host core for
    `cpu_iterations` iterations after the same warm-up."""
    import kutta_amd
    settings = kutta_amd.SolverSettings(viscosity=1.0 / reynolds, momentum_relaxation=relaxation,
                                        pressure_relaxation=1.0, monitor_interval=monitor)
    conditions = kutta_amd.cavity_conditions()
    t = time.perf_counter()
    solver = kutta_amd.SteadySolver(dev, size, size, nz, conditions, settings)
    setup_s = time.perf_counter() - t
    solver.set_graphs(graphs)
    solver.iterate(warmup)
    dev.synchronize()
    waits0 = solver.info()["host_waits"]
    # The first cpu_iterations of the timed window are also timed on their
    # own, the window the CPU baseline covers.
    window = min(cpu_iterations, iterations) if cpu_iterations > 0 else 0
    t0 = time.perf_counter()
    if window:
        solver.iterate(window)
        dev.synchronize()
    t1 = time.perf_counter()
    solver.iterate(iterations - window)
    dev.synchronize()
    t2 = time.perf_counter()
    ms = (t2 - t0) * 1e3 / max(iterations, 1)
    window_ms = (t1 - t0) * 1e3 / window if window else None
    info = solver.info()
    report = solver.report()
    out = {"workload": f"lid-driven cavity {size}x{size}x{nz} hex, Re={reynolds:g}, paired state, "
                       f"SOU + Barth-Jespersen, automatic pseudo time step, AMG-PCG pressure",
           "cells": info["cells"], "ms_per_iteration": round(ms, 4), "timed_iterations": iterations,
           "first_timed_iteration": warmup + 1, "setup_s": round(setup_s, 3),
           "linear_solver_host_waits_per_iteration": round((info["host_waits"] - waits0) / max(iterations, 1), 2),
           "last_pressure_iterations": int(report["pressure_iterations"]),
           "maximum_velocity": report["maximum_velocity"]}
    if window:
        out["ms_per_iteration_cpu_window"] = round(window_ms, 4)
        out["cpu_window"] = f"iterations {warmup + 1}-{warmup + window}"
    out["roofline"] = logical_traffic(dev, solver, ms)
    solver.close()
    if cpu_iterations > 0:
        out["cpu_baseline"] = simplec_cpu_baseline(size, nz, settings, conditions, warmup, cpu_iterations,
                                                   window_ms)
    return out


def simplec_cpu_baseline(size, nz, settings, conditions, warmup, iterations, gpu_ms) -> dict:
    """The oracle (serial restatement of the reference's CPU path) on one host
    core: `iterations` SIMPLEC iterations after the same warm-up."""
    import kutta_amd
    import oracle
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    kinds, wall = kutta_amd.pack_conditions(conditions)
    cpu = oracle.SteadySolver(size, size, nz, kinds, wall, settings.packed())
    cpu.iterate(warmup)
    t0 = time.perf_counter()
    cpu.iterate(iterations)
    ms = (time.perf_counter() - t0) * 1e3 / iterations
    cpu.close()
    return {"value": round(ms, 3), "unit": "ms per SIMPLEC iteration", "cores": 1, "kind": "port",
            "sample": f"{iterations} iterations after {warmup} warm-up iterations of the same cavity",
            "gpu_speedup_same_iterations": round(ms / gpu_ms, 1)}


# The backward-facing step (a k-omega SST channel): air at 20 m/s, 1%
# intensity with turbulent viscosity ratio 10 at the inlet and at outlet
# backflow, h = 0.1 m, 5h upstream, 30h downstream, span 4h between symmetry
# planes, 8h channel above the step, walls graded from a first cell of y+ ~ 1
# (flat-plate skin friction 0.5 m from the inlet) by 1.15.
STEP_SPEED, STEP_RHO, STEP_MU = 20.0, 1.225, 1.7894e-5


def step_case(nz: int):
    import kutta_amd
    from kutta_amd import meshes
    first = meshes.first_cell_height(STEP_SPEED, STEP_MU / STEP_RHO, 0.5)
    mesh = meshes.backward_step(meshes.StepGeometry(), first=first, growth=1.15, largest=0.02, nz=nz)
    turb = kutta_amd.IntensityViscosityRatio(0.01, 10.0)
    cond = {"inlet": kutta_amd.VelocityInlet((STEP_SPEED, 0.0, 0.0), turb),
            "outlet": kutta_amd.PressureOutlet(0.0, None, False, turb),
            "wall": kutta_amd.Wall(), "symmetry": kutta_amd.Symmetry()}
    settings = kutta_amd.SolverSettings(density=STEP_RHO, viscosity=STEP_MU, monitor_interval=20)
    k0 = 1.5 * (0.01 * STEP_SPEED) ** 2
    w0 = STEP_RHO * k0 / (STEP_MU * 10.0)
    return mesh, cond, settings, k0, w0, first


def simplec_step_bench(dev, nz: int, warmup: int, iterations: int, cpu_nz: int, cpu_iterations: int,
                       graphs: bool = True, kcycle: bool = True) -> dict:
    """ms per steady SST SIMPLEC outer iteration on the backward-facing step
    (step_case), nz cells across the span: uniform inflow state, `warmup`
    iterations untimed, then `iterations` timed. The oracle is timed on one
    host core on the same geometry with cpu_nz spanwise cells (fewer cells:
    the serial oracle takes about a minute per iteration at full size).
    With kcycle the same iterations are timed again with the pressure AMG
    rebuilt as the K-cycle (set_pressure_options; no reference counterpart)."""
    import kutta_amd
    mesh, cond, settings, k0, w0, first = step_case(nz)
    t = time.perf_counter()
    solver = kutta_amd.SteadySolver.from_mesh(dev, mesh, cond, settings, kutta_amd.SstSettings())
    setup_s = time.perf_counter() - t
    solver.set_graphs(graphs)
    t = time.perf_counter()
    solver.initialize_fields((STEP_SPEED, 0.0, 0.0), 0.0, k0, w0)
    solver.iterate(warmup)
    dev.synchronize()
    warm_s = time.perf_counter() - t
    waits0 = solver.info()["host_waits"]
    t0 = time.perf_counter()
    solver.iterate(iterations)
    dev.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / max(iterations, 1)
    info = solver.info()
    report = solver.report(monitor=True)
    out = {"workload": (f"SST k-omega SIMPLEC, backward-facing step h=0.1 m (5h up, 30h down, 8h channel, "
                        f"span 4h symmetry), {mesh.n_cells} hex cells, first cell {first:.3g} m growth 1.15, "
                        f"air 20 m/s, TI 1%, viscosity ratio 10, batched momentum, AMG-PCG V(2,2) pressure"),
           "cells": info["cells"], "faces": mesh.n_faces, "ms_per_iteration": round(ms, 3),
           "timed_iterations": iterations, "first_timed_iteration": warmup + 1, "setup_s": round(setup_s, 2),
           "initialize_and_warmup_s": round(warm_s, 2), "batched_momentum": solver.set_batched_momentum(True),
           "linear_solver_host_waits_per_iteration": round((info["host_waits"] - waits0) / max(iterations, 1), 2),
           "last_pressure_iterations": int(report["pressure_iterations"]),
           "last_k_iterations": int(report["k_iterations"]), "last_omega_iterations": int(report["omega_iterations"]),
           "residuals": {k: report[k] for k in ("continuity_scaled", "u_scaled", "v_scaled", "k_scaled",
                                                "omega_scaled")},
           "maximum_velocity": report["maximum_velocity"],
           "pressure_levels": solver.pressure_levels(), "pressure_plan": solver.pressure_plan()}
    out["roofline"] = logical_traffic(dev, solver, ms, 2)
    if kcycle:
        # The same iterations with the pressure AMG rebuilt as the K-cycle
        # (KCYCLE, 3 aggregation passes as the default hierarchy): fields
        # re-initialised to the uniform inflow state, same warm-up.
        from kutta_amd import AmgCycle, AmgOptions, OperatorStructure
        solver.set_graphs(graphs)
        solver.initialize_fields((STEP_SPEED, 0.0, 0.0), 0.0, k0, w0)
        solver.set_pressure_options(AmgOptions(structure=OperatorStructure.SymmetricMMatrix, cycle=AmgCycle.K,
                                               **KCYCLE))
        solver.iterate(warmup)
        dev.synchronize()
        t0 = time.perf_counter()
        solver.iterate(iterations)
        dev.synchronize()
        kms = (time.perf_counter() - t0) * 1e3 / max(iterations, 1)
        kreport = solver.report(monitor=True)
        out["pressure_kcycle"] = {
            "options": f"K-cycle V(1,1), 4 coarsest sweeps, 5 levels, inner corrections on levels of "
                       f"{KCYCLE['k_min_rows']} rows or more",
            "ms_per_iteration": round(kms, 3), "speedup": round(ms / kms, 3),
            "last_pressure_iterations": int(kreport["pressure_iterations"]),
            "residuals": {k: kreport[k] for k in ("continuity_scaled", "u_scaled", "v_scaled", "k_scaled",
                                                  "omega_scaled")},
            "pressure_plan": solver.pressure_plan()}
        print(f"[bench] SST step with K-cycle pressure: {kms:.3f} ms per iteration "
              f"({out['pressure_kcycle']['last_pressure_iterations']} pressure iterations last)",
              file=sys.stderr, flush=True)
    solver.close()
    if cpu_iterations > 0:
        import numpy as np

        import oracle
        cmesh, ccond, csettings, _, _, _ = step_case(cpu_nz)
        try:
            os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
        except (AttributeError, OSError):
            pass
        cpu = oracle.SteadySolver.from_mesh(cmesh, kutta_amd.pack_mesh_conditions(cmesh, ccond),
                                            csettings.packed(), kutta_amd.SstSettings().packed())
        n = cmesh.n_cells
        cpu.initialize_fields(np.full(n, STEP_SPEED), np.zeros(n), np.zeros(n), np.zeros(n), np.full(n, k0),
                              np.full(n, w0))
        cpu.iterate(warmup)
        t0 = time.perf_counter()
        cpu.iterate(cpu_iterations)
        cms = (time.perf_counter() - t0) * 1e3 / cpu_iterations
        cpu.close()
        scaled = cms * mesh.n_cells / n
        out["cpu_baseline"] = {
            "value": round(cms, 1), "unit": "ms per SIMPLEC iteration", "cores": 1, "kind": "port",
            "sample": (f"{cpu_iterations} iterations after {warmup} warm-up iterations of the same step with "
                       f"{cpu_nz} spanwise cells ({n} cells)"),
            "per_cell_scaled_to_full_mesh_ms": round(scaled, 1), "gpu_speedup_per_cell": round(scaled / ms, 1)}
    return out


def trace_summary(db: str, line_json: str | None, top: int) -> int:
    """Markdown summary of a rocprofv3 --kernel-trace database (rocpd SQLite):
    busy time per kernel and grid size (one row per AMG level for the level
    kernels), and for the l1-Jacobi smoother (thread- or wavefront-per-row
    form) the algorithmic bytes 20n + 8m
    of each level (rows n, nonzeros m from the bench line's "levels"),
    GB/s and the fraction of the 8 TB/s HBM peak. With an --fv-kernels line
    the finite-volume kernels get their algorithmic bytes from it."""
    import re
    import sqlite3
    rows = sqlite3.connect(db).execute("select name, grid_x, duration from kernels").fetchall()

    def short(name):
        n = name.replace("(anonymous namespace)::", "").replace("void ", "")
        n = re.sub(r"^kh::", "", n)
        # The wavefront-per-row kernels share one template: name them by form.
        if n.startswith("k_rowwave<"):
            if "Prolonged" in n:
                return "k_prolong_smooth_rowwave"
            return "k_csr_residual_rowwave" if "RowWaveOut)1" in n else "k_amg_smooth_rowwave"
        m = re.match(r"k_each<(?:kh::)?(?:fv::)?([A-Za-z_0-9:]+)", n)
        return "fv " + m.group(1) if m else re.sub(r"[<(].*", "", n)

    agg = {}
    for name, grid, ns in rows:
        a = agg.setdefault((short(name), grid), [0, 0.0])
        a[0] += 1
        a[1] += ns
    total = sum(v[1] for v in agg.values())
    levels, fv_bytes = [], {}
    if line_json:
        line = json.loads(Path(line_json).read_text())
        levels = line["config"].get("levels") or line["config"].get("pressure_levels", [])
        # An --fv-kernels line: algorithmic bytes per kernel, keyed by the
        # functor's name (fv_paired_limiter -> FvPairedLimiter).
        for k in line.get("kernels", []):
            fv_bytes["fv " + "".join(w.capitalize() for w in k["kernel"].split("_"))] = k["bytes"]
    print(f"busy {total / 1e6:.2f} ms over {len(rows)} dispatches\n")
    print("| kernel | grid | launches | total ms | avg us | % busy | level rows | bytes/launch | GB/s | frac |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
    for (name, grid), (count, ns) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        avg_us = ns / count / 1e3
        extra = ["", "", "", ""]
        if name in ("k_amg_smooth", "k_amg_smooth_rowwave"):
            # thread per row: 256 rows per workgroup; wavefront per row: 4 rows
            # per 256-thread workgroup. Levels sharing a grid size stay unnamed.
            per_group = 256 if name == "k_amg_smooth" else 4
            match = [(n_rows, nnz) for n_rows, nnz in levels if grid == -(-n_rows // per_group) * 256]
            if len(match) == 1:
                n_rows, nnz = match[0]
                b = 20 * n_rows + 8 * nnz
                gbs = b / (avg_us * 1e-6) / 1e9
                extra = [str(n_rows), str(b), f"{gbs:.0f}", f"{gbs / HBM_PEAK_GBS:.3f}"]
            elif match:
                extra = ["/".join(str(r) for r, _ in match), "", "", ""]
        elif name in fv_bytes:
            b = fv_bytes[name]
            gbs = b / (avg_us * 1e-6) / 1e9
            extra = ["", str(b), f"{gbs:.0f}", f"{gbs / HBM_PEAK_GBS:.3f}"]
        print(f"| {name} | {grid} | {count} | {ns / 1e6:.2f} | {avg_us:.2f} | {100 * ns / total:.1f} | "
              + " | ".join(extra) + " |")
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=55)
    ap.add_argument("--sweeps", type=int, default=5)
    ap.add_argument("--tail", type=int, default=-1,
                    help="first AMG level of the one-workgroup coarse-grid tail (-1: chosen by the "
                         "autotuner, 0 off)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every PCG iteration kernel by kernel instead of replaying a captured HIP "
                         "graph (identical iterates; for kernel tracers)")
    ap.add_argument("--fv-kernels", action="store_true",
                    help="only time each finite-volume kernel on the size^3 mesh")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only time device copies of BYTES (counter calibration)")
    ap.add_argument("--simplec-size", type=int, default=128,
                    help="cells per side of the 2-D cavity timed beside the solve (0 skips it)")
    ap.add_argument("--simplec-iterations", type=int, default=500)
    ap.add_argument("--simplec-box", type=int, default=0, metavar="N",
                    help="only time SIMPLEC on the N^3 lid-driven box (the reference's FV benchmark)")
    ap.add_argument("--direct-rows", type=int, default=8192,
                    help="mixed variants: the first coarse level of at most this many rows is solved by "
                         "its FP64 dense inverse")
    ap.add_argument("--variants", default="all",
                    help="comma-separated solver variants to time beside the plain solve "
                         "(pcg_mixed, pcg_mixed_bf16, pcg_mixed_lazy, fgmres_plain, fgmres_mixed, fgmres_learned, "
                         "fgmres_mixed_learned), "
                         "'all' or 'none'")
    ap.add_argument("--learned-train", type=int, default=None, metavar="STEPS",
                    help="exactly STEPS Adam steps instead of --learned-epochs epochs (0 skips the learned variants)")
    ap.add_argument("--learned-harvest", type=int, default=None, metavar="K",
                    help="harvest from K right-hand sides instead of until --learned-pairs pairs are held")
    ap.add_argument("--learned-epochs", type=int, default=2,
                    help="training epochs of the learned coarse correction (0 skips the learned variants)")
    ap.add_argument("--learned-pairs", type=int, default=1000,
                    help="(residual, correction) pairs the FGMRES solver harvests for training, 90%% trained "
                         "on, 10%% held out (20000 is the specified scale; its exact targets take about six "
                         "minutes, so the default line uses fewer)")
    ap.add_argument("--learned-level", type=int, default=2, help="AMG level the network works on")
    ap.add_argument("--learned-max-unexplained", type=float, default=0.5,
                    help="largest unexplained fraction of the training targets for which the "
                         "network is applied (larger: apply it regardless)")
    ap.add_argument("--trace-summary", metavar="DB", help="summarize a rocprofv3 kernel-trace database and exit")
    ap.add_argument("--trace-line", metavar="JSON", help="the bench line of the traced run (AMG level sizes)")
    ap.add_argument("--trace-top", type=int, default=40)
    ap.add_argument("--large-size", type=int, default=110,
                    help="cells per side of the box on which the plain path and the K-cycle variant are also "
                         "timed (110: 7,986,000 cells; 0 skips it)")
    ap.add_argument("--step-nz", type=int, default=80,
                    help="spanwise cells of the SST step in the default line (0 skips it)")
    ap.add_argument("--simplec-step", type=int, default=0, metavar="NZ",
                    help="only time SST SIMPLEC on the backward-facing step with NZ spanwise cells (80: 4.06M)")
    ap.add_argument("--no-step-kcycle", action="store_true",
                    help="skip the SST step's second timing with the K-cycle pressure AMG")
    ap.add_argument("--simplec-only", action="store_true",
                    help="only time SIMPLEC on the 2-D cavity (--simplec-size, --simplec-iterations)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe, max(args.steps, 1))
    if args.fv_kernels:
        return fv_kernel_bench(args.size, max(args.steps, 1))
    if args.trace_summary:
        return trace_summary(args.trace_summary, args.trace_line, args.trace_top)
    if args.simplec_step:
        import kutta_amd
        dev = kutta_amd.Device(0)
        line = simplec_step_bench(dev, args.simplec_step, max(args.warmup, 1), max(args.steps, 1), 2,
                                  0 if args.no_cpu_baseline else 2, not args.no_graphs, not args.no_step_kcycle)
        line = {"metric": "ms/SIMPLEC iteration", "value": line["ms_per_iteration"], "unit": "ms",
                "n_gpus": 1, "higher_is_better": False, "dtype": "f32", "data": "synthetic", "config": line}
        print(json.dumps(line), flush=True)
        dev.close()
        return 0
    if args.simplec_box or args.simplec_only:
        import kutta_amd
        dev = kutta_amd.Device(0)
        if args.simplec_box:
            # bench/bench_fv.cpp:128-143: default settings except momentum
            # relaxation 1, iterate(5) untimed, then timed iterations.
            n = args.simplec_box
            line = simplec_bench(dev, n, n, 100.0, 1.0, 20, 5, max(args.steps, 1), 0, not args.no_graphs)
        else:
            line = simplec_bench(dev, args.simplec_size, 1, 1000.0, 1.0, 100, 5,
                                 max(args.simplec_iterations - 5, 1), 0 if args.no_cpu_baseline else 20,
                                 not args.no_graphs)
        line = {"metric": "ms/SIMPLEC iteration", "value": line["ms_per_iteration"], "unit": "ms",
                "n_gpus": 1, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
                "config": line}
        print(json.dumps(line), flush=True)
        dev.close()
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not a run)", file=sys.stderr)
        return 2

    import numpy as np

    import kutta_amd
    from kutta_amd import (AmgCycle, AmgOptions, Kernel, OperatorStructure, PressureProblem, SolveControl,
                           SolverKind)

    dev = kutta_amd.Device(0)
    t = time.perf_counter()
    problem = PressureProblem(dev, args.size, args.size, args.size, tet=True, amplitude=0.2,
                              seed=SEED, reference=-2)
    setup_mesh_s = time.perf_counter() - t
    options = AmgOptions(pre_sweeps=args.sweeps, post_sweeps=args.sweeps, aggregation_passes=2,
                         structure=OperatorStructure.SymmetricMMatrix)
    t = time.perf_counter()
    solver = kutta_amd.make_amg_solver(SolverKind.AmgPcg, dev, problem.host, options)
    setup_amg_s = time.perf_counter() - t
    tail_level = solver.set_tail(args.tail)
    solver.set_graphs(not args.no_graphs)
    t = time.perf_counter()
    solver.refresh(problem)
    dev.synchronize()
    plain_refresh_ms = (time.perf_counter() - t) * 1e3
    if args.tail < 0:  # automatic: the autotuner times each admissible tail level (identical iterates)
        t = time.perf_counter()
        tail_level = solver.autotune_tail(problem, 5)
        autotune_s = time.perf_counter() - t
    else:
        autotune_s = 0.0
    b_host = rhs_for(problem.n_cells, problem.reference)
    b = dev.upload(b_host)
    x = dev.alloc(problem.n_cells)
    control = SolveControl(rel_tol=1e-6, abs_tol=0.0, max_iter=2000)
    print(f"[bench] {problem.n_cells} cells, nnz {problem.nnz}, {solver.describe()}, tail level {tail_level}", file=sys.stderr,
          flush=True)

    for _ in range(args.warmup):
        x.fill(np.float32(0))
        solver.solve(problem, x, b, control)
    dev.synchronize()

    n, m = problem.n_cells, problem.nnz
    stats = []
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x.fill(np.float32(0))
        stats.append(solver.solve(problem, x, b, control))
    dev.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed * 1e3 / args.steps

    # Fine-level smoothing sweeps are timed on the device stream by HIP events
    # attached to their own dispatch packets (hipExtLaunchKernelGGL), so the
    # average is the kernel's execution time, as a kernel trace reports it.
    # Graph-replayed launches carry no events, so these solves launch kernel
    # by kernel (the same kernels and iterates).
    solver.set_graphs(False)
    dev.profile_reset()
    dev.profile(Kernel.amg_smooth, True)
    dev.profile_filter(Kernel.amg_smooth, n)
    for _ in range(max(args.steps // 2, 2)):
        x.fill(np.float32(0))
        solver.solve(problem, x, b, control)
    dev.synchronize()
    launches, kernel_ms = dev.profile_read(Kernel.amg_smooth)
    dev.profile(Kernel.amg_smooth, False)
    solver.set_graphs(not args.no_graphs)
    # The same sweep as a batch of back-to-back launches with one pair of HIP
    # events around the batch (kh_time_launches): execution plus the gap
    # between launches, the rate a sequence of sweeps sustains.
    repeats = 50
    inverse, invalid, swept = dev.alloc(n), dev.alloc(n), dev.alloc(n)
    guard = dev.upload(np.zeros(3, np.int32))
    dev.launch(Kernel.krylov_diagonal, n, problem.row_ptr, problem.diag, problem.val, inverse, invalid, 1)
    smooth_args = (problem.row_ptr, problem.col, problem.val, x, b, inverse, swept, guard)
    dev.launch(Kernel.amg_smooth, n, *smooth_args)
    batch_us = dev.time_launches(Kernel.amg_smooth, n, *smooth_args, repeats=repeats) * 1e3 / repeats
    st = stats[-1]
    if not st.converged:
        print(f"[bench] warning: solve did not converge ({st})", file=sys.stderr)

    # amg_smooth reads x, b, l1-inverse, row_ptr and writes next (5 x 4n) and
    # gathers col + val (8m) once: 20n + 8m bytes per launch.
    bytes_per_launch = 20 * n + 8 * m
    traffic = measured_traffic() if args.size == 55 else None
    avg_us = kernel_ms * 1e3 / max(launches, 1)
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9 if launches else None
    roofline = {
        "kernel": "amg_smooth (fine level)",
        "bound": "hbm",
        "achieved": round(achieved, 1) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
        "traffic": traffic,
        "bytes_per_launch": bytes_per_launch,
        "avg_launch_us": round(avg_us, 3),
        "launches": launches,
        "batch_avg_us": round(batch_us, 3),
        "batch_frac": round(bytes_per_launch / (batch_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
    }

    # The same solve with the additive solver variants (no reference
    # counterpart): the mixed-precision V-cycle (first coarse level of at most
    # --direct-rows rows solved by an FP64 dense inverse; with pcg_mixed_bf16
    # also the coarse levels' values in bfloat16) under PCG, flexible
    # GMRES(30) with the plain and the mixed V-cycle, and the K-cycle variant
    # (KCYCLE) alone and with the dense level. x is compared with the plain
    # FP32 solution and with the FP64 solution, all within the 1e-6 relative
    # residual. Each variant's refresh (the operator update every SIMPLEC
    # outer iteration pays) is timed beside its solve.
    variants = {}
    x_plain = x.download()

    def gpu_solve(rhs):
        x.fill(np.float32(0))
        solver.solve(problem, x, dev.upload(rhs), control)
        return x.download()

    t = time.perf_counter()
    x64, x64_residual, x64_steps = fp64_reference(problem, b_host, x_plain, gpu_solve)
    fp64 = {"method": "iterative refinement: FP64 residuals on the host, FP32 AMG-PCG corrections on the GPU",
            "relative_residual": x64_residual, "refinement_steps": x64_steps,
            "seconds": round(time.perf_counter() - t, 2), "plain_error": fp64_error(x_plain, x64)}
    wanted = (None if args.variants == "all" else
              set() if args.variants == "none" else set(args.variants.split(",")))
    dr = args.direct_rows
    k_options = AmgOptions(cycle=AmgCycle.K, aggregation_passes=2, structure=OperatorStructure.SymmetricMMatrix,
                           **KCYCLE)
    for label, kind, opts, direct, bf16 in (("pcg_mixed", SolverKind.AmgPcg, options, dr, False),
                                            ("pcg_mixed_bf16", SolverKind.AmgPcg, options, dr, True),
                                            ("fgmres_plain", SolverKind.AmgFgmres, options, 0, False),
                                            ("fgmres_mixed", SolverKind.AmgFgmres, options, dr, False),
                                            ("pcg_kcycle", SolverKind.AmgPcg, k_options, 0, False),
                                            ("pcg_kcycle_mixed", SolverKind.AmgPcg, k_options, dr, False)):
        if wanted is not None and label not in wanted:
            continue
        _, vst, out = time_solver(dev, problem, opts, kind, b, x, control, args.steps, args.warmup,
                                  not args.no_graphs, direct, bf16, tail=-1 if opts is k_options else args.tail)
        xv = x.download()
        out["speedup_vs_plain_pcg"] = round(ms / out["ms"], 3)
        out["speedup_refresh_plus_solve"] = round((ms + plain_refresh_ms) / out["refresh_plus_solve_ms"], 3)
        out["fp64_error"] = fp64_error(xv, x64)
        out["max_abs_diff_vs_plain_over_max_abs_x"] = float(np.abs(xv.astype(np.float64) - x_plain).max() /
                                                            np.abs(x_plain).max())
        variants[label] = out
        print(f"[bench] {label}: {out['ms']:.3f} ms, {vst.iterations} iterations", file=sys.stderr, flush=True)

    # The mixed path's lazy dense refresh over a drifting operator.
    if wanted is None or "pcg_mixed_lazy" in wanted:
        variants["pcg_mixed_lazy"] = lazy_dense_sequence(dev, problem, options, b, x, control, dr,
                                                         not args.no_graphs)
        v = variants["pcg_mixed_lazy"]
        print(f"[bench] pcg_mixed_lazy: {v['mixed_lazy']['refresh_plus_solve_ms']:.3f} ms refresh + solve "
              f"({v['mixed_lazy']['dense_inversions']} inversions, {v['mixed_lazy']['dense_kept']} kept); "
              f"eager {v['mixed_eager']['refresh_plus_solve_ms']:.3f}, plain {v['plain']['refresh_plus_solve_ms']:.3f}",
              file=sys.stderr, flush=True)

    # The learned coarse correction inside FGMRES(30) (learned_variant):
    # harvest, targets and training are setup, reported beside the timed ms.
    if args.learned_epochs > 0 and args.learned_train != 0:
        for label, direct in (("fgmres_learned", 0), ("fgmres_mixed_learned", args.direct_rows)):
            if wanted is not None and label not in wanted:
                continue
            if wanted is None and label == "fgmres_mixed_learned":
                continue  # the C4-scale run is one variant; name it to run this one
            variants[label] = learned_variant(dev, problem, options, b, x, x_plain, x64, control, ms,
                                              direct, args)
            v = variants[label]
            print(f"[bench] {label}: {v['ms']:.3f} ms, {v['iterations']} iterations, {v['pairs']} pairs, "
                  f"unexplained {v['unexplained_by_epoch']} held-out {v['unexplained_held_out']:.3f}, "
                  f"gated {v['gated']}/{v['solves_in_learned_mode']}; applied regardless: "
                  f"{v['forced']['ms']:.3f} ms, {v['forced']['iterations']} iterations", file=sys.stderr, flush=True)

    simplec = None
    if args.simplec_size:
        # Configuration 1: 128x128x1 cavity, Re=1000, relaxation 1, monitors
        # every 100, 500 iterations; ms averaged over iterations 6-500.
        simplec = simplec_bench(dev, args.simplec_size, 1, 1000.0, 1.0, 100, 5,
                                max(args.simplec_iterations - 5, 1), 0 if args.no_cpu_baseline else 20,
                                not args.no_graphs)

    simplec_sst = None
    if args.step_nz:
        # Configuration 3: SST on the backward-facing step (4.06M hex cells at
        # 80 spanwise cells), 3 untimed iterations from the uniform inflow
        # state, then 10 timed.
        simplec_sst = simplec_step_bench(dev, args.step_nz, 3, 10, 2, 0 if args.no_cpu_baseline else 2,
                                         not args.no_graphs, not args.no_step_kcycle)

    large = None
    if args.large_size:
        large = large_mesh_comparison(dev, args.large_size, max(min(args.steps, 5), 1), 1, not args.no_graphs, args)

    baseline = None
    x_gpu = x_plain
    if not args.no_cpu_baseline:
        baseline, x_cpu = cpu_baseline(problem, b_host, args.sweeps, st)
        baseline["x_bitwise_equal_gpu"] = bool(np.array_equal(x_cpu.view(np.uint32), x_gpu.view(np.uint32)))

    line = {
        "metric": METRIC,
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"AMG-PCG pressure solve to 1e-6: {args.size}^3 Freudenthal tet box "
                         f"({n} cells), +-0.2h seeded perturbation, origin cell pinned, seeded "
                         f"zero-mean Gaussian RHS, aggregation AMG 2 passes V({args.sweeps},{args.sweeps}), "
                         f"plain FP32 path"),
            "cells": n,
            "nnz": m,
            "hierarchy": solver.describe(),
            "plan": solver.plan(),
            "levels": solver.levels(),
            "coarse_tail_level": tail_level,
            "tail_autotune_s": round(autotune_s, 3),
            "iterations": st.iterations,
            "relative_residual": st.final_residual / st.initial_residual,
            "converged": st.converged,
            "refresh_ms": round(plain_refresh_ms, 2),
            "setup_s": {"mesh_and_assembly": round(setup_mesh_s, 3), "amg": round(setup_amg_s, 3),
                        **{k: round(v, 3) for k, v in problem.setup_seconds.items()}},
            "device": dev.name,
        },
        "roofline": roofline,
        "cpu_baseline": baseline,
        "fp64_reference": fp64,
        "solver_variants": variants,
        "large_mesh": large,
        "simplec": simplec,
        "simplec_sst": simplec_sst,
    }
    print(json.dumps(line), flush=True)
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
