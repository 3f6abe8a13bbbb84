"""Admission logic of the native request loop (csrc/serve.cpp), without a GPU.

Jobs here have no solver: their setup callback sleeps and returns NULL, so
only the scheduler runs. Checks: the admitted footprints never exceed the
budget, admission is first fit in arrival order (a small job behind a large
one that does not fit yet starts first), every job completes, oversize jobs
are refused."""
from __future__ import annotations

import ctypes as C
import threading
import time


SETUP_FUNC = C.CFUNCTYPE(C.c_void_p, C.c_void_p)
DONE_FUNC = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)


def test_admission_by_footprint_first_fit(gpu_lib):
    lib = gpu_lib
    srv = lib.IBMServer_new(100.0, 3)
    lock = threading.Lock()
    running = {"now": 0.0, "peak": 0.0}
    order = []
    sizes = [60.0, 50.0, 30.0, 20.0, 90.0, 10.0]
    keep = []

    def make(i, size):
        def setup(_u):
            with lock:
                running["now"] += size
                running["peak"] = max(running["peak"], running["now"])
                order.append(i)
            time.sleep(0.15)
            with lock:
                running["now"] -= size
            return None

        def done(_h, _u):
            pass
        cs, cd = SETUP_FUNC(setup), DONE_FUNC(done)
        keep.append((cs, cd))
        return cs, cd

    ids = []
    for i, size in enumerate(sizes):
        cs, cd = make(i, size)
        ids.append(lib.IBMServer_submit(srv, size, 0, C.cast(cs, C.c_void_p), C.cast(cd, C.c_void_p), None))
    assert ids == list(range(len(sizes)))
    cs, cd = make(99, 101.0)
    assert lib.IBMServer_submit(srv, 101.0, 0, C.cast(cs, C.c_void_p), C.cast(cd, C.c_void_p), None) == -1
    lib.IBMServer_wait(srv)
    t = (C.c_double * 3)()
    admitted = {}
    for j in ids:
        assert lib.IBMServer_job_times(srv, j, t) == 2
        assert 0 <= t[0] <= t[1] <= t[2]
        admitted[j] = t[1]
    peak = lib.IBMServer_peak_bytes(srv)
    lib.IBMServer_destroy(srv)
    assert running["peak"] <= 100.0 and peak <= 100.0
    # 60 runs; 50 does not fit beside it, 30 does (first fit), then 10.
    # Jobs admitted together start on different workers in any order, so
    # the admission times (taken in arrival order) are checked, not `order`.
    assert admitted[0] <= admitted[2] <= admitted[5] < admitted[1]
    assert sorted(order) == list(range(len(sizes)))


def _configured(I, n=(8, 8, 8), precision=0, solver=None, periodic_x=None):
    """An unassembled cavity-like solver; periodic_x: None, "both" or "west"."""
    import numpy as np
    s = I.Solver()
    nx, ny, nz = n
    s.set_grid(np.linspace(0, 1, nx + 1), np.linspace(0, 1, ny + 1), np.linspace(0, 1, nz + 1))
    s.set_params(100.0, 0.01)
    s.set_bc(I.DIR_NORTH | I.DIR_EAST | I.DIR_WEST | I.DIR_SOUTH | I.DIR_DOWN | I.DIR_UP, I.BC_STATIONARY_WALL)
    if periodic_x == "both":
        s.set_bc(I.DIR_WEST | I.DIR_EAST, I.BC_ALL_PERIODIC)
    elif periodic_x == "west":
        s.set_bc(I.DIR_WEST, I.BC_ALL_PERIODIC)
    s.set_linear_solver(solver if solver is not None else I.SOLVER_BiCGSTAB, I.PRECOND_AMG, 1e-10)
    s.set_precision(precision)
    return s


def test_failed_jobs_fail_alone(I, gpu_lib, capfd):
    """Jobs whose input IBMSolver_assemble rejects fail with the reference's
    message while the server (and this process) goes on: the error is a job
    state, its budget is released, and a job without a grid fails the same
    way. All four fail before any device work, so no GPU is needed."""
    from importlib import import_module
    serving = import_module("2020researchproject_amd.serving")
    srv = serving.Server(1000.0, 2, gpu_lib._name)
    bad = [
        (_configured(I, periodic_x="west"), "Inconsistent periodic boundary condition"),
        (_configured(I, n=(9, 8, 8), periodic_x="both"), "periodic x axis with an odd cell count (9)"),
        (_configured(I, precision=I.PREC_MIXED, solver=I.SOLVER_GMRES), "IBM3D_PREC_MIXED supports only"),
        (I.Solver(), "IBMSolver_set_grid must precede IBMSolver_assemble"),
    ]
    res = [srv.submit_solver(s, None, 400.0, 3) for s, _ in bad]
    srv.wait()
    for r, (_, msg) in zip(res, bad):
        srv.collect(r)
        assert r.state == 3 and msg in r.error, (r.state, r.error)
        assert 0 <= r.submit <= r.admit <= r.finish
    assert srv.failed_jobs() == 4
    # 400 + 400 fit the budget of 1000 together, a third did not; whether
    # two of these instant failures overlapped at all depends on timing
    assert srv.peak_bytes() in (400.0, 800.0)
    srv.close()
    out = capfd.readouterr()
    assert "Inconsistent periodic boundary condition" in out.out
    assert "odd cell count (9)" in out.err


def test_native_submission_refusals(I, gpu_lib):
    """IBMServer_submit_solver refuses a NULL solver and a footprint above the
    budget (-1, the solver stays the caller's); IBMServer_job_result of an
    unknown job is -1."""
    import ctypes as C
    srv = gpu_lib.IBMServer_new(100.0, 1)
    assert gpu_lib.IBMServer_submit_solver(srv, None, 10.0, 1, None, None) == -1
    s = _configured(I)
    assert gpu_lib.IBMServer_submit_solver(srv, s.h, 101.0, 1, None, None) == -1
    res = I.ServeResult()
    assert gpu_lib.IBMServer_job_result(srv, 0, C.byref(res)) == -1
    assert gpu_lib.IBMServer_failed_jobs(srv) == 0
    gpu_lib.IBMServer_destroy(srv)
    s.close()  # still the caller's: destroyed here, unassembled


def test_pressure_mg_bits_before_assembly(I, gpu_lib):
    """IBMSolver_get_pressure_mg_bits reports 0 until the solver is assembled
    (host state only, no GPU needed)."""
    s = _configured(I, precision=I.PREC_MIXED_KRYLOV)
    assert s.pressure_mg_bits() == 0
    s.close()
