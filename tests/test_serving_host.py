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
