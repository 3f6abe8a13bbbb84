"""Host-side logic: workload builders and the seeded perturbation (CPU)."""
from __future__ import annotations

import numpy as np


def test_uniform01_scalar_matches_vectorised(cases):
    xc = np.array([-0.03125, 0.03125, 0.5, 0.96875])
    yc = np.array([0.15625, 0.71875])
    zc = np.array([0.40625, 0.09375, 1.03125])
    g = cases.uniform01_grid(2020, 2, xc, yc, zc)
    for a, x in enumerate(xc):
        for b, y in enumerate(yc):
            for c, z in enumerate(zc):
                assert g[a, b, c] == cases.uniform01(2020, 2, x, y, z)
    assert 0.0 <= g.min() and g.max() < 1.0


def test_uniform01_is_uniform(cases):
    x = np.linspace(0.001, 0.999, 64)
    g = cases.uniform01_grid(5, 1, x, x, x[:16]).ravel()
    assert abs(g.mean() - 0.5) < 0.01 and abs(g.var() - 1 / 12) < 0.005


def test_sphere_faces_clustered_and_monotone(cases):
    xf, yf, zf = cases.sphere_faces()
    for f, a, b, n in ((xf, -5, 15, 256), (yf, -5, 5, 128), (zf, -5, 5, 128)):
        d = np.diff(f)
        assert len(f) == n + 1 and f[0] == a and f[-1] == b and (d > 0).all()
        assert abs(d.min() - cases.SPHERE_DMIN) < 2e-3          # smallest cells at the body
        i0 = np.searchsorted(f, 0.0)
        assert d[i0 - 1] < 1.1 * d.min() and d.max() < 0.2
        assert (np.abs(d[1:] / d[:-1] - 1) < 0.2).all()          # cell-to-cell growth below 20%


def test_cell_centres_match_assemble_rule(cases):
    f = np.array([0.0, 0.1, 0.3, 0.6, 1.0])
    c = cases.cell_centres(f)
    np.testing.assert_allclose(c, [-0.2, -0.05, 0.05, 0.2, 0.45, 0.8, 1.2, 1.55])  # mirrored widths


def test_cavity_init_array_equals_init_func(I, cases, new_oracle):
    a, b = new_oracle(), new_oracle()
    cases.setup_cavity(a, n=6, init="array")
    cases.setup_cavity(b, n=6, init="func")
    for name in ("U1", "U2", "U3", "P"):
        assert np.array_equal(a.get(name), b.get(name))


def test_bench_step_roofline_arithmetic():
    """The whole-step figure of the bench line on hand-made timer entries:
    kernels a and c state bytes (4 launches of 1 GB, 2 of 0.5 GB = 5 GB over 2
    steps in 4 ms of kernel time), b states none (1 ms); steps of 10 ms."""
    import bench
    stats = [{"name": "a", "ms": 2.0, "count": 4, "bytes": 1e9},
             {"name": "b", "ms": 1.0, "count": 2, "bytes": 0.0},
             {"name": "c", "ms": 2.0, "count": 2, "bytes": 0.5e9}]
    r = bench.step_roofline(stats, 2, 1000, np.array([10.0, 10.0, 12.0]))
    assert r["bytes_per_cell_step"] == 2.5e6
    assert r["achieved_kernels"] == 1250.0 and r["frac_kernels"] == 0.1562
    assert r["achieved_step"] == 250.0 and r["frac_step"] == 0.0312
    assert r["kernel_time_with_bytes"] == 0.8 and r["peak"] == 8000.0
    assert bench.step_roofline([], 2, 1000, np.array([10.0])) is None
    assert bench.step_roofline(stats, 0, 1000, np.array([10.0])) is None
