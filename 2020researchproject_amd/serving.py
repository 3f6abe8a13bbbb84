"""Serving stream of independent cases over the native request loop.

The request loop itself is C++ (csrc/serve.cpp, include/ibm3d_serve.h):
admission by device-memory footprint (first fit in arrival order), one
worker thread and one HIP stream per running case, cases overlapping on the
GPU. The submitting thread configures each case's solver (IBMSolver_new and
the IBMSolver_set_* calls, host work only) and hands it over with
IBMServer_submit_solver; the worker assembles it, sets the initial flow, runs
the steps and records the results in native code, so no Python runs on the
workers and a case error fails only its job. This module builds the seeded
stream of the mixed-size workload (lid-driven cavities and analytic-sphere
cases arriving as a Poisson process), submits each case at its arrival time,
and reports time-to-solution percentiles and aggregate cell-steps/s. There is
no counterpart in the reference, whose drivers run one case per MPI job.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import cases
from . import ibm3d as I

GRIDS = ((32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 64, 64))


@dataclass
class CaseSpec:
    kind: str                  # "cavity" or "sphere"
    grid: tuple
    Re: float
    steps: int
    arrival: float             # seconds after the stream starts
    tol: float = 1e-10
    precision: int = I.PREC_MIXED_KRYLOV

    @property
    def cells(self) -> int:
        return int(np.prod(self.grid))


def make_stream(n: int = 64, mean_interarrival: float = 0.050, seed: int = 5, steps=(100, 400),
                grids=GRIDS, tol: float = 1e-10, precision: int = I.PREC_MIXED_KRYLOV) -> list[CaseSpec]:
    """n cases: exponential inter-arrival times, cavity or sphere with
    probability 1/2 each, grid uniform over `grids`, Re ~ U[100, 1000],
    steps ~ U{steps[0]..steps[1]} (all drawn from one seeded generator)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(mean_interarrival, n))
    t -= t[0]
    out = []
    for c in range(n):
        kind = "cavity" if rng.random() < 0.5 else "sphere"
        grid = grids[rng.integers(len(grids))]
        Re = float(rng.uniform(100.0, 1000.0))
        st = int(rng.integers(steps[0], steps[1] + 1))
        out.append(CaseSpec(kind, tuple(grid), Re, st, float(t[c]), tol, precision))
    return out


def build_case(spec: CaseSpec, lib=None) -> I.Solver:
    """The case built and initialised in this process (the reference run a
    served case is compared with)."""
    s = I.Solver(lib)
    nx, ny, nz = spec.grid
    if spec.kind == "cavity":
        cases.setup_cavity(s, Re=spec.Re, tol=spec.tol, precision=spec.precision, shape=(nx, ny, nz))
    else:
        cases.setup_sphere(s, nx, ny, nz, Re=spec.Re, tol=spec.tol, precision=spec.precision)
    return s


def configure_case(spec: CaseSpec, lib=None) -> tuple[I.Solver, I.ServeInit]:
    """The case's solver configured but not assembled, and its initial flow,
    for IBMServer_submit_solver."""
    s = I.Solver(lib)
    nx, ny, nz = spec.grid
    if spec.kind == "cavity":
        init = cases.setup_cavity(s, Re=spec.Re, tol=spec.tol, precision=spec.precision, shape=(nx, ny, nz),
                                  assemble=False)
    elif spec.kind == "sphere":
        init = cases.setup_sphere(s, nx, ny, nz, Re=spec.Re, tol=spec.tol, precision=spec.precision,
                                  assemble=False)
    else:
        raise ValueError(f"unknown case kind {spec.kind!r}")
    return s, init


@dataclass
class JobResult:
    spec: CaseSpec | None
    job: int
    state: int = 0             # 0 queued, 1 running, 2 done, 3 failed
    submit: float = -1.0
    admit: float = -1.0
    finish: float = -1.0
    iters: list = field(default_factory=list)
    max_div: float = float("nan")
    fields: dict | None = None
    device_bytes: float = 0.0
    error: str | None = None


class Server:
    """Python handle of the native request loop."""

    def __init__(self, budget_bytes: float, max_active: int, lib=None):
        self.lib = I.load(lib)
        self.h = self.lib.IBMServer_new(float(budget_bytes), int(max_active))
        self._keep: dict[int, tuple] = {}

    def estimate_bytes(self, spec: CaseSpec) -> float:
        nx, ny, nz = spec.grid
        return self.lib.IBMSolver_estimate_device_bytes(nx, ny, nz, spec.precision, 0)

    def submit_solver(self, s: I.Solver, init: I.ServeInit | None, bytes_: float, steps: int,
                      keep_fields: bool = False, res: JobResult | None = None) -> JobResult:
        """Hands the configured, unassembled solver `s` to the server (which
        owns it from here on) as a job of `steps` steps."""
        res = res or JobResult(None, -1)
        shape = (s.Nx + 4, s.Ny + 4, s.Nz + 4) if keep_fields else None
        arrays = [np.zeros(shape) for _ in range(4)] if keep_fields else []
        ptrs = (C.c_void_p * 4)(*[a.ctypes.data for a in arrays]) if keep_fields else None
        init_ref = C.byref(init) if init is not None else None
        h = s.detach()
        res.job = self.lib.IBMServer_submit_solver(self.h, h, float(bytes_), int(steps), init_ref, ptrs)
        if res.job < 0:
            self.lib.IBMSolver_destroy(h)
            raise ValueError(f"case does not fit the budget ({bytes_:.3g} bytes)")
        # the Solver object stays referenced: it holds the ctypes callbacks
        # (function-valued boundary data) the worker's native code may call
        self._keep[res.job] = (res, arrays, ptrs, init, s)
        return res

    def submit(self, spec: CaseSpec, keep_fields: bool = False) -> JobResult:
        s, init = configure_case(spec, self.lib)
        return self.submit_solver(s, init, self.estimate_bytes(spec), spec.steps, keep_fields,
                                  JobResult(spec, -1))

    def wait(self):
        self.lib.IBMServer_wait(self.h)

    def result(self, job: int) -> I.ServeResult:
        out = I.ServeResult()
        self.lib.IBMServer_job_result(self.h, job, C.byref(out))
        return out

    def collect(self, res: JobResult) -> JobResult:
        """Copies the finished job's outcome into res."""
        r = self.result(res.job)
        res.state = r.state
        res.submit, res.admit, res.finish = tuple(r.t)
        res.iters = list(r.iters)
        res.max_div = r.max_div
        res.device_bytes = r.device_bytes
        res.error = r.error.decode() or None
        arrays = self._keep[res.job][1] if res.job in self._keep else []
        if arrays and r.state == 2:
            res.fields = dict(zip(("U1", "U2", "U3", "P"), arrays))
        return res

    def times(self, job: int):
        out = (C.c_double * 3)()
        state = self.lib.IBMServer_job_times(self.h, job, out)
        return state, tuple(out)

    def failed_jobs(self) -> int:
        return self.lib.IBMServer_failed_jobs(self.h)

    def peak_bytes(self) -> float:
        return self.lib.IBMServer_peak_bytes(self.h)

    def close(self):
        if self.h:
            self.lib.IBMServer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_stream(specs: list[CaseSpec], budget_bytes: float, max_active: int = 4, keep_fields: bool = False,
               lib=None, allow_failures: bool = False) -> dict:
    """Submits every case at its arrival time (relative to the first) and
    waits for all of them. Returns per-job results and the summary:
    time-to-solution = completion - submission, aggregate Mcell-steps/s =
    sum(cells * steps) / (last completion - first submission), the cell-steps
    of the cases that finished. A failed case raises unless allow_failures."""
    srv = Server(budget_bytes, max_active, lib)
    results: list[JobResult] = []
    t0 = time.perf_counter()
    for spec in sorted(specs, key=lambda c: c.arrival):
        delay = spec.arrival - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        results.append(srv.submit(spec, keep_fields))
    srv.wait()
    for r in results:
        srv.collect(r)
    peak = srv.peak_bytes()
    srv.close()
    errors = [r.error for r in results if r.state != 2]
    if errors and not allow_failures:
        raise RuntimeError(f"{len(errors)} case(s) failed: {errors[:3]}")
    tts = np.array([r.finish - r.submit for r in results])
    span = max(r.finish for r in results) - min(r.submit for r in results)
    work = sum(r.spec.cells * r.spec.steps for r in results if r.state == 2)
    return {
        "results": results,
        "p50_s": float(np.percentile(tts, 50)),
        "p99_s": float(np.percentile(tts, 99)),
        "makespan_s": float(span),
        "mcell_steps_per_s": work / span / 1e6,
        "peak_admitted_bytes": peak,
        "budget_bytes": budget_bytes,
        "max_active": max_active,
        "cases": len(results),
    }
