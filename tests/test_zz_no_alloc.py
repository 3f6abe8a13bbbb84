"""Polyhedron_cpt_noalloc (no device allocation; every buffer in the
library's fixed device scratch) against Polyhedron_cpt, through the C-ABI:
bit-for-bit equal distances on two calls in a row, the second on a smaller
grid (the scratch keeps the first call's contents), and a request larger
than the scratch refused on the host with nothing written."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def _poly(lib):
    from test_gpu_parity import read_poly
    return read_poly(lib, os.path.join(GOLDEN, "sphere.stl"))


def _cpt(fn, poly, x, y, z, maxd=15.0):
    f = np.full((len(x), len(y), len(z)), -7.0)
    rc = fn(C.c_void_p(poly), len(x), len(y), len(z), x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
            z.ctypes.data_as(_dp), f.ctypes.data_as(_dp), C.c_double(maxd))
    return rc, f


def test_cpt_noalloc_matches_cpt_twice(gpu_lib):
    lib = gpu_lib
    lib.Polyhedron_cpt.restype = None
    lib.Polyhedron_cpt_noalloc.restype = C.c_int
    for fn in (lib.Polyhedron_cpt, lib.Polyhedron_cpt_noalloc):
        fn.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double]
    poly = _poly(lib)
    try:
        grids = [(np.linspace(-1.2, 1.3, 61), np.linspace(-1.1, 1.0, 47), np.linspace(-0.9, 1.4, 53)),
                 (np.linspace(-0.8, 0.9, 23), np.linspace(-1.0, 0.7, 19), np.linspace(-0.6, 0.8, 17))]
        for x, y, z in grids:
            print(f"grid {len(x)}x{len(y)}x{len(z)}", flush=True)
            _, ref = _cpt(lib.Polyhedron_cpt, poly, x, y, z)
            rc, got = _cpt(lib.Polyhedron_cpt_noalloc, poly, x, y, z)
            assert rc == 0
            assert np.isfinite(ref).all()
            assert np.array_equal(got, ref)
            assert (ref < 0).any() and (ref > 0).any()
        # more points than the scratch holds: refused before any device work
        big = np.linspace(0.0, 1.0, 256)
        rc, f = _cpt(lib.Polyhedron_cpt_noalloc, poly, big, big, big)
        assert rc == -1 and (f == -7.0).all()
    finally:
        lib.Polyhedron_destroy(poly)
