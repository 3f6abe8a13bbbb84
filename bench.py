#!/usr/bin/env python3
"""Benchmark of the fractional-step IBM time step on one MI355X.

Workload: flow past a sphere (phi = |x| - 0.5), domain [-5,15]x[-5,5]^2,
256x128x128 cells cosine-stretched toward the body, Re = 300, dt = 0.002,
uniform inflow u = 1, convective outflow, free-slip lateral walls, linear
solves to relative residual 1e-9 (which keeps max|div U| <= 1e-8 for both
precisions). A step is one IBMSolver_iterate time step (AB2 convection,
three Crank-Nicolson Helmholtz solves, pressure Poisson solve, projection,
boundary and immersed-boundary ghost updates, and the drag/lift reduction
of the step's fields). Inputs are resident in HBM when the timed region
starts.

The headline value is the mixed path in Mcell-steps/s: FP64 BiCGSTAB whose
multigrid preconditioner runs in FP32 (IBM3D_PREC_MIXED_KRYLOV). The plain
FP64 path runs the same steps for the speed-up. After the timed regions both
GPU paths and the FP64 CPU oracle run the first steps of the same case from
the same initial field; their differences (rel-L2 and Linf of u and p,
max|div U|, drag and lift) are the line's error statement. Prints one JSON
line.
"""
from __future__ import annotations

import argparse
import importlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG = importlib.import_module("2020researchproject_amd")
I, cases = PKG.ibm3d, PKG.cases

METRIC = "Mcell-steps/s & p99 latency: FP32-MG+FP64 refine vs FP64, same div tol"
HBM_PEAK_GBS = 8000.0
ORACLE_LIB = os.path.join(ROOT, "oracle", "libibm3d_oracle.so")
HALF_AREA = 0.5 * math.pi * 0.25  # (1/2) U^2 A_ref, U = 1, A_ref = pi r^2 of the r = 0.5 sphere


def coefficients(force):
    """Drag and lift coefficients (Cd, Cl_y, Cl_z) of a body force."""
    return [float(f) / HALF_AREA for f in force]


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def run_gpu(prec: int, nx: int, ny: int, nz: int, steps: int, warmup: int, tol: float, prof_steps: int,
            graphs: bool = True, outflow: str = "convective"):
    """Set up the sphere case, run `warmup` untimed steps, time `steps`
    steps (no per-kernel instrumentation: multigrid cycles replay as HIP
    graphs), then run `prof_steps` more steps with every kernel bracketed by
    HIP events on the solver's stream for the per-kernel table."""
    s = I.Solver()
    t0 = time.perf_counter()
    s.set_graph_replay(graphs)
    cases.setup_sphere(s, nx, ny, nz, tol=tol, precision=prec, outflow=outflow)
    log(f"[bench] setup prec={prec} {nx}x{ny}x{nz}: {time.perf_counter() - t0:.2f} s, "
        f"ghost cells {len(s.ghost_cells())}")
    if warmup:
        s.iterate(warmup)
        log(f"[bench] warmup {warmup} steps done (its {s.last_iters().tolist()})")
    s.set_force_history(steps)  # drag / lift of every timed step, kept on the device
    t0 = time.perf_counter()
    s.iterate(steps)
    wall = time.perf_counter() - t0
    step_ms = s.step_times()
    force_hist = s.force_history()
    s.set_force_history(0)
    res = {
        "wall": wall, "step_ms": step_ms, "max_div": s.max_div(),
        "iters": s.last_iters().tolist(), "norms": s.last_norms().tolist(),
        "u": [s.get(n)[2:-2, 2:-2, 2:-2] for n in ("U1", "U2", "U3")],
        "p": s.get("P")[2:-2, 2:-2, 2:-2],
        "force": s.body_force().tolist(),
        "coef_hist": [coefficients(f) for f in force_hist],
        "mg_levels": ["x".join(str(v) for v in lv) for lv in s.mg_levels()],
    }
    log(f"[bench] prec={prec} timed {steps} steps: {wall:.3f} s, max|div U| {res['max_div']:.3e}, "
        f"its {res['iters']}")
    res["stats"] = []
    if prof_steps:
        s.enable_timing(2)
        s.reset_timing()
        s.iterate(prof_steps)
        res["graph_stats"] = s.kernel_stats()
        s.enable_timing(1)
        s.reset_timing()
        s.iterate(prof_steps)
        res["stats"] = s.kernel_stats()
        s.enable_timing(0)
        log(f"[bench] prec={prec} instrumented {prof_steps} steps done")
    s.close()
    return res


def rel_l2(a, b):
    m = np.isfinite(a) & np.isfinite(b)
    return float(np.linalg.norm(a[m] - b[m]) / max(np.linalg.norm(b[m]), 1e-300))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))


def fields(s):
    u = [s.get(n)[2:-2, 2:-2, 2:-2] for n in ("U1", "U2", "U3")]
    return u, s.get("P")[2:-2, 2:-2, 2:-2]


def errors(u, p, ref_u, ref_p):
    """rel-L2 and Linf (relative to the largest reference value) of the
    velocity as a vector field and of p, over the cells where both are finite
    (NaN marks solid cells in both)."""
    m = np.isfinite(ref_p) & np.isfinite(p)
    for a, b in zip(u, ref_u):
        m &= np.isfinite(a) & np.isfinite(b)
    du = np.sqrt(sum((a[m] - b[m]) ** 2 for a, b in zip(u, ref_u)))
    su = np.sqrt(sum(b[m] ** 2 for b in ref_u))
    dp = p[m] - ref_p[m]
    return {"rel_l2_u": float(np.linalg.norm(du) / max(np.linalg.norm(su), 1e-300)),
            "linf_u": float(du.max() / max(su.max(), 1e-300)),
            "rel_l2_p": float(np.linalg.norm(dp) / max(np.linalg.norm(ref_p[m]), 1e-300)),
            "linf_p": float(np.abs(dp).max() / max(np.abs(ref_p[m]).max(), 1e-300))}


def vs_oracle(nx, ny, nz, tol, steps, outflow, mixed_mode):
    """Both GPU paths and the FP64 CPU oracle (all host threads) run `steps`
    steps of the benchmark case from the same initial field. Returns the
    error statement and the oracle's timing (the CPU baseline on the full
    workload)."""
    o = I.Solver(ORACLE_LIB)
    cases.setup_sphere(o, nx, ny, nz, tol=tol, outflow=outflow)
    log(f"[bench] oracle {nx}x{ny}x{nz} assembled; {steps} steps on {host_threads()} host threads")
    t0 = time.perf_counter()
    o.iterate(steps)
    t_oracle = time.perf_counter() - t0
    log(f"[bench] oracle {steps} steps: {t_oracle:.1f} s, its {o.last_iters().tolist()}")
    ref_u, ref_p = fields(o)
    ref = {"max_div": o.max_div(), "coef": coefficients(o.body_force()), "iters_last_step": o.last_iters().tolist()}
    ref_flags, ref_ghosts = o.flags(), o.ghost_cells()
    o.close()
    out = {"steps": steps, "grid": f"{nx}x{ny}x{nz}", "oracle": ref}
    for tag, prec in (("mixed", mixed_mode), ("fp64", I.PREC_FP64)):
        g = I.Solver()
        cases.setup_sphere(g, nx, ny, nz, tol=tol, precision=prec, outflow=outflow)
        if tag == "mixed":
            out["flags_bit_exact"] = bool(np.array_equal(g.flags(), ref_flags))
            out["ghost_cells_bit_exact"] = bool(np.array_equal(g.ghost_cells(), ref_ghosts))
        g.iterate(steps)
        u, p = fields(g)
        e = errors(u, p, ref_u, ref_p)
        e["max_div"] = g.max_div()
        e["coef"] = coefficients(g.body_force())
        e["coef_abs_err"] = max(abs(a - b) for a, b in zip(e["coef"], ref["coef"]))
        out[tag] = e
        g.close()
        log(f"[bench] {tag} vs oracle: {e}")
    cb = {"value": nx * ny * nz * steps / t_oracle / 1e6, "unit": "Mcell-steps/s", "cores": host_threads(),
          "kind": "port", "cpu_model": cpu_model(),
          "sample": f"the first {steps} steps of the same sphere case at full size ({nx}x{ny}x{nz}), FP64 "
                    f"oracle (OpenMP, block-ILU(0) BiCGSTAB pressure), {t_oracle:.1f} s"}
    return out, cb


def cavity_full(tol: float = 1e-10, steps: int = 200):
    """The 32^3 lid-driven cavity (Re 100, dt 0.005, seeded perturbation) run
    in full on the CPU oracle (all host threads) and on the GPU in FP64; the
    oracle's rate and the GPU's deviation from it. Also the oracle on one
    thread for a 20-step sample of the same case."""
    o = I.Solver(ORACLE_LIB)
    cases.setup_cavity(o, n=32, tol=tol)
    t0 = time.perf_counter()
    o.iterate(steps)
    t_full = time.perf_counter() - t0
    ref_u, ref_p = fields(o)
    o.close()
    g = I.Solver()
    cases.setup_cavity(g, n=32, tol=tol)
    g.iterate(steps)
    u, p = fields(g)
    g.close()
    e = errors(u, p, ref_u, ref_p)
    # the pressure of the cavity is defined up to the pinned cell; both pin p'(0,0,0)
    import ctypes
    gomp = ctypes.CDLL("libgomp.so.1")
    n1 = 20
    o = I.Solver(ORACLE_LIB)
    cases.setup_cavity(o, n=32, tol=tol)
    gomp.omp_set_num_threads(1)
    try:
        t0 = time.perf_counter()
        o.iterate(n1)
        t_one = time.perf_counter() - t0
    finally:
        gomp.omp_set_num_threads(host_threads())
    o.close()
    cells = 32 ** 3
    log(f"[bench] cavity 32^3: oracle {steps} steps {t_full:.1f} s, one thread {n1} steps {t_one:.1f} s; "
        f"GPU FP64 vs oracle {e}")
    return {"workload": f"cavity 32^3, Re 100, dt 0.005, {steps} steps, tol {tol:g}",
            "value": round(cells * steps / t_full / 1e6, 4), "seconds": round(t_full, 2),
            "gpu_fp64_vs_oracle": e,
            "one_thread": {"value": round(cells * n1 / t_one / 1e6, 4), "sample": f"the same case, {n1} steps"}}


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes recorded in profiles/traffic.json (each counter scaled
    by its reading on a plain device copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f).get("kernels", {}).get(kernel)
    return None if k is None else k["traffic_bytes"]


def roofline(stats):
    """Roofline entry of the streaming kernel with the largest total time
    (kernels with a stated algorithmic byte count per launch)."""
    cand = [t for t in stats if t["bytes"] > 0 and t["count"] > 0]
    if not cand:
        return None
    dom = max(cand, key=lambda t: t["ms"])
    avg_ms = dom["ms"] / dom["count"]
    achieved = dom["bytes"] / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom["name"]),
            "kernel": dom["name"], "avg_us": round(avg_ms * 1e3, 2), "bytes_per_launch": round(dom["bytes"]),
            "launches": dom["count"], "traffic_source": "profiles/traffic.json"}


def step_roofline(stats, prof_steps: int, cells: int, step_ms):
    """The whole step against the HBM roofline: the algorithmic bytes of
    every timed kernel that states them (each distinct array element once per
    launch), summed over the instrumented steps and given per cell-step; that
    sum over those kernels' summed duration ("kernels": every kernel at its
    own pace) and, per step, over the median time of the timed steps
    ("step": launch gaps and host waits included)."""
    if not stats or prof_steps <= 0:
        return None
    nbytes = sum(t["bytes"] * t["count"] for t in stats if t["bytes"] > 0)
    t_bytes = sum(t["ms"] for t in stats if t["bytes"] > 0) * 1e-3
    t_all = sum(t["ms"] for t in stats) * 1e-3
    per_step = nbytes / prof_steps
    wall = float(np.median(step_ms)) * 1e-3
    if t_bytes <= 0 or wall <= 0:
        return None
    kern, step = nbytes / t_bytes / 1e9, per_step / wall / 1e9
    return {"bytes_per_cell_step": round(per_step / cells, 1), "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "achieved_kernels": round(kern, 1), "frac_kernels": round(kern / HBM_PEAK_GBS, 4),
            "achieved_step": round(step, 1), "frac_step": round(step / HBM_PEAK_GBS, 4),
            "kernel_time_with_bytes": round(t_bytes / t_all, 3)}


def top_kernels(stats, n: int = 8):
    """[name, share of timed GPU time, avg us, algorithmic GB/s] of the
    kernels with the largest total time in the instrumented steps."""
    tot = sum(t["ms"] for t in stats) or 1.0
    out = []
    for t in sorted(stats, key=lambda t: -t["ms"])[:n]:
        avg = t["ms"] / max(t["count"], 1)
        gbs = t["bytes"] / (avg * 1e-3) / 1e9 if avg > 0 and t["bytes"] else None
        out.append([t["name"], round(t["ms"] / tot, 3), round(avg * 1e3, 1), None if gbs is None else round(gbs, 1)])
    return out


def copy_calibration(mib: float) -> None:
    """One plain device-to-device copy of `mib` MiB (torch's copy kernel):
    its rocprofv3 FETCH_SIZE / WRITE_SIZE readings against the known byte
    count give the scale of those counters."""
    import torch
    n = int(mib * (1 << 20)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: {n * 4} bytes read and written")
    del a, b


def serving_stream(args) -> dict:
    """Serving stream: n independent cases (lid-driven cavity or analytic
    sphere with probability 1/2; grids 32^3, 48^3, 64^3, 96^3, 128x64x64; Re
    ~ U[100, 1000]; 100-400 steps; seed 5) arriving with exponential gaps of
    mean 50 ms, mixed precision, pressure tolerance 1e-10, through the native
    request loop (each case configured here, assembled, initialised and run
    by a native worker). Aggregate Mcell-steps/s = all cases' cell-steps over
    first arrival to last completion; p50/p99 time-to-solution = completion -
    arrival; mean_active = summed service times / makespan (the cases' overlap)."""
    serving = importlib.import_module("2020researchproject_amd.serving")
    specs = serving.make_stream(n=args.serving_cases, seed=5)
    budget = args.serving_budget_gib * (1 << 30)
    t0 = time.perf_counter()
    res = serving.run_stream(specs, budget_bytes=budget, max_active=args.serving_slots, allow_failures=True)
    wall = time.perf_counter() - t0
    busy = sum(r.finish - r.admit for r in res["results"])
    return {
        "value": round(res["mcell_steps_per_s"], 2), "unit": "Mcell-steps/s",
        "workload": f"{len(specs)} cases, Poisson arrivals (mean 50 ms), cavity/sphere, "
                    "grids 32^3..128x64x64, Re 100-1000, 100-400 steps, tol 1e-10, FP64 BiCGSTAB + FP32 multigrid",
        "slots": args.serving_slots, "budget_gib": args.serving_budget_gib,
        "cases": len(specs), "failed": sum(r.state != 2 for r in res["results"]),
        "p50_s": round(res["p50_s"], 3), "p99_s": round(res["p99_s"], 3),
        "makespan_s": round(res["makespan_s"], 3), "wall_s": round(wall, 3),
        "mean_active": round(busy / res["makespan_s"], 2),
        "cell_steps": int(sum(c.cells * c.steps for c in specs)),
        "peak_admitted_gib": round(res["peak_admitted_bytes"] / (1 << 30), 3),
    }


def run_serving(args) -> int:
    """The serving stream alone, as its own JSON line."""
    sv = serving_stream(args)
    out = {
        "metric": "serving stream: aggregate Mcell-steps/s, p50/p99 time-to-solution",
        "value": sv["value"], "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": sv["workload"], "slots": sv["slots"], "budget_gib": sv["budget_gib"]},
    }
    out.update({k: v for k, v in sv.items() if k not in ("value", "unit", "workload", "slots", "budget_gib")})
    print(json.dumps(out), flush=True)
    return 0


def print_stats(tag: str, res) -> None:
    for key in ("graph_stats", "stats"):
        for t in sorted(res.get(key, []), key=lambda t: -t["ms"]):
            avg = t["ms"] / max(t["count"], 1)
            gbs = t["bytes"] / (avg * 1e-3) / 1e9 if avg > 0 and t["bytes"] else 0
            log(f"[{tag} {key}] {t['name']:18s} total {t['ms']:9.2f} ms  n {t['count']:6d}  "
                f"avg {avg * 1e3:9.1f} us  {gbs:8.1f} GB/s")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200,
                    help="timed steps (200: p50/p99 of the per-step times over 200 samples)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=str, default="256x128x128")
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--mixed-mode", type=int, default=2, choices=(1, 2),
                    help="1: FP64 defect correction around FP32 BiCGSTAB; 2: FP64 BiCGSTAB, FP32 multigrid")
    ap.add_argument("--copy-calib-mb", type=float, default=0.0,
                    help="also run one plain device copy of this many MiB (calibrates rocprofv3 byte counters)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--prof-steps", type=int, default=2, help="instrumented steps for the per-kernel table")
    ap.add_argument("--stats", action="store_true", help="print per-kernel timing to stderr")
    ap.add_argument("--no-graphs", action="store_true", help="launch multigrid cycles kernel by kernel")
    ap.add_argument("--outflow", choices=("convective", "pressure"), default="convective",
                    help="East boundary: convective outflow (the workload as stated) or the reference's "
                         "pressure outlet")
    ap.add_argument("--oracle-steps", type=int, default=2,
                    help="steps of the full-size comparison with the FP64 CPU oracle (0: skip)")
    ap.add_argument("--serving", action="store_true",
                    help="run only the serving stream (64 cavity/sphere cases, Poisson arrivals)")
    ap.add_argument("--no-serving", action="store_true", help="leave the serving stream out of the bench line")
    ap.add_argument("--serving-cases", type=int, default=64)
    ap.add_argument("--serving-slots", type=int, default=12,
                    help="cases running at once (measured best of 4, 8, 12, 16 on the 64-case stream)")
    ap.add_argument("--serving-budget-gib", type=float, default=64.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU (one GPU holds the whole domain); --gpus {args.gpus} "
              "is not supported", file=sys.stderr)
        return 2
    if args.serving:
        return run_serving(args)
    nx, ny, nz = (int(v) for v in args.grid.split("x"))
    cells = float(nx * ny * nz)

    if args.copy_calib_mb > 0:
        copy_calibration(args.copy_calib_mb)
    mixed = run_gpu(args.mixed_mode, nx, ny, nz, args.steps, args.warmup, args.tol, args.prof_steps,
                    not args.no_graphs, args.outflow)
    value = cells * args.steps / mixed["wall"] / 1e6
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(mixed["wall"] / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"sphere {nx}x{ny}x{nz} cosine-stretched, Re 300, dt 0.002, uniform inflow, "
                               f"{args.outflow} outflow, slip lateral walls, tol {args.tol:g}",
                   "cells": int(cells),
                   "precision": ("momentum: FP32 red-black sweeps in FP64 refinement; pressure: FP64 BiCGSTAB, "
                                 "FP32 multigrid preconditioner" if args.mixed_mode == 2 else
                                 "FP64 defect correction around FP32 multigrid-BiCGSTAB"),
                   "mg_levels": mixed["mg_levels"]},
        "p50_ms": round(float(np.percentile(mixed["step_ms"], 50)), 3),
        "p99_ms": round(float(np.percentile(mixed["step_ms"], 99)), 3),
        "max_div": mixed["max_div"],
        "iters_last_step": mixed["iters"],
        # drag and lift coefficients (Cd, Cl_y, Cl_z) after each timed step
        "cd_history": [round(c[0], 6) for c in mixed["coef_hist"]],
        "cl_history": [[float(f"{c[1]:.6g}"), float(f"{c[2]:.6g}")] for c in mixed["coef_hist"]],
    }
    rf = roofline(mixed["stats"])
    if rf:
        out["roofline"] = rf
    out["top_kernels"] = top_kernels(mixed["stats"])
    sr = step_roofline(mixed["stats"], args.prof_steps, cells, mixed["step_ms"])
    if sr:
        out["step_roofline"] = sr
    if args.stats:
        print_stats("mixed", mixed)
    if not args.no_fp64:
        f64 = run_gpu(I.PREC_FP64, nx, ny, nz, args.steps, args.warmup, args.tol, args.prof_steps,
                      not args.no_graphs, args.outflow)
        v64 = cells * args.steps / f64["wall"] / 1e6
        out["fp64_value"] = round(v64, 2)
        out["speedup_vs_fp64"] = round(value / v64, 3)
        out["fp64_max_div"] = f64["max_div"]
        out["fp64_p99_ms"] = round(float(np.percentile(f64["step_ms"], 99)), 3)
        out["rel_l2_u_mixed_vs_fp64"] = max(rel_l2(a, b) for a, b in zip(mixed["u"], f64["u"]))
        out["fp64_iters_last_step"] = f64["iters"]
        out["cd_abs_diff_mixed_vs_fp64"] = max(abs(a[0] - b[0]) for a, b in zip(mixed["coef_hist"], f64["coef_hist"]))
        sr = step_roofline(f64["stats"], args.prof_steps, cells, f64["step_ms"])
        if sr:
            out["fp64_step_roofline"] = sr
        if args.stats:
            print_stats("fp64", f64)
    if not args.no_serving:
        log("[bench] serving stream")
        out["serving"] = serving_stream(args)
    if args.oracle_steps > 0 and os.path.exists(ORACLE_LIB):
        acc, cb = vs_oracle(nx, ny, nz, args.tol, args.oracle_steps, args.outflow, args.mixed_mode)
        out["vs_oracle"] = acc
        if not args.no_cpu:
            cb["c1_cavity_full"] = cavity_full()
            out["cpu_baseline"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in cb.items()}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
