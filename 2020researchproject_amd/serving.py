"""Serving stream of independent cases over the native request loop.

The request loop itself is C++ (csrc/serve.cpp, include/ibm3d_serve.h):
admission by device-memory footprint (first fit in arrival order), one
worker thread and one HIP stream per running case, cases overlapping on the
GPU. This module builds the seeded stream of the mixed-size workload (lid-
driven cavities and analytic-sphere cases arriving as a Poisson process),
submits each case at its arrival time, and reports time-to-solution
percentiles and aggregate cell-steps/s. There is no counterpart in the
reference, whose drivers run one case per MPI job.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import cases
from . import ibm3d as I

SETUP_FUNC = C.CFUNCTYPE(C.c_void_p, C.c_void_p)
DONE_FUNC = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)

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
    s = I.Solver(lib)
    nx, ny, nz = spec.grid
    if spec.kind == "cavity":
        cases.setup_cavity(s, Re=spec.Re, tol=spec.tol, precision=spec.precision, shape=(nx, ny, nz))
    else:
        cases.setup_sphere(s, nx, ny, nz, Re=spec.Re, tol=spec.tol, precision=spec.precision)
    return s


@dataclass
class JobResult:
    spec: CaseSpec
    job: int
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

    def submit(self, spec: CaseSpec, keep_fields: bool = False) -> JobResult:
        res = JobResult(spec, -1)
        box: dict = {}

        def _setup(_user):
            try:
                s = build_case(spec, self.lib)
                res.device_bytes = s.device_bytes()
                box["solver"] = s
                return s.detach()
            except Exception as e:  # reported with the job
                res.error = repr(e)
                return None

        def _done(h, _user):
            s = box.get("solver")
            if s is None or not h:
                return
            s.h = h  # borrowed until this callback returns; the server destroys it
            try:
                res.iters = s.last_iters().tolist()
                res.max_div = s.max_div()
                if keep_fields:
                    res.fields = {n: s.get(n) for n in ("U1", "U2", "U3", "P")}
            except Exception as e:
                res.error = repr(e)
            finally:
                s.h = None

        cs, cd = SETUP_FUNC(_setup), DONE_FUNC(_done)
        res.job = self.lib.IBMServer_submit(self.h, self.estimate_bytes(spec), spec.steps,
                                            C.cast(cs, C.c_void_p), C.cast(cd, C.c_void_p), None)
        if res.job < 0:
            raise ValueError(f"case does not fit the budget: {spec}")
        self._keep[res.job] = (cs, cd, box)
        return res

    def wait(self):
        self.lib.IBMServer_wait(self.h)

    def times(self, job: int):
        out = (C.c_double * 3)()
        state = self.lib.IBMServer_job_times(self.h, job, out)
        return state, tuple(out)

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
               lib=None) -> dict:
    """Submits every case at its arrival time (relative to the first) and
    waits for all of them. Returns per-job results and the summary:
    time-to-solution = completion - submission, aggregate Mcell-steps/s =
    sum(cells * steps) / (last completion - first submission)."""
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
        _, (r.submit, r.admit, r.finish) = srv.times(r.job)
    peak = srv.peak_bytes()
    srv.close()
    errors = [r.error for r in results if r.error]
    if errors:
        raise RuntimeError(f"{len(errors)} case(s) failed: {errors[:3]}")
    tts = np.array([r.finish - r.submit for r in results])
    span = max(r.finish for r in results) - min(r.submit for r in results)
    work = sum(r.spec.cells * r.spec.steps for r in results)
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
