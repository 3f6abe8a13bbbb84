"""Pins of the FP64 CPU oracle (CPU only).

The reference ships no tests or golden vectors for this path, so every oracle
function is pinned here by what the mathematics fixes: closed forms on small
grids, exactness on linear fields, operator identities, brute force on tiny
inputs, and analytic solutions. Each check is chosen so that a plausible
restatement mistake (dropped term, wrong index, transposed weight, wrong
spacing) fails it.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _p(a):
    return a.ctypes.data_as(_dp)


def bind_oracle_hooks(lib):
    lib.oracle_apply_matrix.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    lib.oracle_matrix_row.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _dp]
    lib.oracle_matrix_row.restype = C.c_int
    lib.oracle_get_kcoef.argtypes = [C.c_void_p, _dp]
    lib.oracle_get_coords.argtypes = [C.c_void_p] + [_dp] * 6
    for n in ("oracle_update_outer", "oracle_update_ghost", "oracle_calc_N"):
        getattr(lib, n).argtypes = [C.c_void_p]
    return lib


def coords(s):
    nx, ny, nz = s.Nx, s.Ny, s.Nz
    arr = [np.zeros(n + 4) for n in (nx, ny, nz, nx, ny, nz)]
    s.lib.oracle_get_coords(s.h, *[_p(a) for a in arr])
    return arr  # xc, yc, zc, dx, dy, dz


def kcoef(s):
    nx, ny, nz = s.Nx, s.Ny, s.Nz
    out = np.zeros(2 * (nx + 4) + 2 * (ny + 4) + 2 * (nz + 4))
    s.lib.oracle_get_kcoef(s.h, _p(out))
    o = np.cumsum([0, nx + 4, nx + 4, ny + 4, ny + 4, nz + 4, nz + 4])
    return [out[o[a]:o[a + 1]] for a in range(6)]


def box(I, s, xf, yf, zf, Re=2.0, dt=4.0, bc=None, tol=1e-12):
    s.set_grid(xf, yf, zf)
    s.set_params(Re, dt)
    if bc is None:
        s.set_bc(I.DIR_EAST | I.DIR_WEST | I.DIR_NORTH | I.DIR_SOUTH | I.DIR_DOWN | I.DIR_UP,
                 I.BC_STATIONARY_WALL)
    else:
        bc(s)
    s.set_obstacle(None)
    s.set_linear_solver(I.SOLVER_BiCGSTAB, I.PRECOND_AMG, tol)
    s.assemble()
    return s


# ---------------------------------------------------------------------------
# a13: grid metrics and K coefficients  <- src/ibm3d_setup.c:536-632
# ---------------------------------------------------------------------------
def test_kcoef_closed_form_nonuniform(I, new_oracle, oracle_lib):
    bind_oracle_hooks(oracle_lib)
    s = box(I, new_oracle(), [0, 1, 3, 6, 10], [0, 1, 2, 3], [0, 2, 4, 6])
    kW, kE, kS, kN, kD, kU = kcoef(s)
    # dt / (2 Re) = 1; dx = 1,2,3,4; xc = .5,2,4.5,8; mirrored ghosts -.5, 12
    np.testing.assert_allclose(kW[2:6], [1 / 1, 1 / 3, 1 / 7.5, 1 / 14], rtol=1e-15)
    np.testing.assert_allclose(kE[2:6], [1 / 1.5, 1 / 5, 1 / 10.5, 1 / 16], rtol=1e-15)
    np.testing.assert_allclose(kS[2:5], [1, 1, 1], rtol=1e-15)
    np.testing.assert_allclose(kU[2:5], [1 / 4, 1 / 4, 1 / 4], rtol=1e-15)
    assert kW[0] == kW[1] == kW[6] == kW[7] == 0  # outer entries stay zero


def test_ghost_widths_periodic(I, new_oracle, oracle_lib):
    bind_oracle_hooks(oracle_lib)

    def per(s):
        s.set_bc(I.DIR_EAST | I.DIR_WEST, I.BC_ALL_PERIODIC)
        s.set_bc(I.DIR_NORTH | I.DIR_SOUTH | I.DIR_DOWN | I.DIR_UP, I.BC_STATIONARY_WALL)
    s = box(I, new_oracle(), [0, 1, 3, 6, 10], [0, 1, 2, 3], [0, 1, 2, 3], bc=per)
    xc, yc, zc, dx, dy, dz = coords(s)
    np.testing.assert_allclose(dx, [3, 4, 1, 2, 3, 4, 1, 2])   # wrapped widths
    np.testing.assert_allclose(xc[:2], [-5.5, -2.0])           # 0.5 - (4+1)/2, -2 - (3+4)/2
    np.testing.assert_allclose(xc[6:], [10.5, 12.0])
    kW, kE = kcoef(s)[:2]
    assert kW[2] == pytest.approx(1 / 2.5, rel=1e-15)


# ---------------------------------------------------------------------------
# a11: mirror-point trilinear stencil  <- src/utils.c:111-171
# ---------------------------------------------------------------------------
def sphere_solver(I, s, cases, n=(24, 20, 20)):
    return cases.setup_sphere(s, *n, tol=1e-10)


def test_ghost_interp_reproduces_mirror_point(I, new_oracle, cases):
    s = sphere_solver(I, new_oracle(), cases)
    xc, yc, zc, *_ = coords(bind_oracle_hooks(s.lib) and s)
    lv = s.get("LVSET")
    ghosts = s.ghost_cells()
    assert len(ghosts) > 20
    idx = np.zeros((8, 3), dtype=np.int32)
    w = np.zeros(8)
    for (i, j, k) in ghosts:
        s.lib.IBMSolver_ghost_interp(s.h, int(i), int(j), int(k), idx.ctypes.data_as(_ip), _p(w))
        L = lambda a, b, c: lv[a + 2, b + 2, c + 2]
        n = np.array([(L(i + 1, j, k) - L(i - 1, j, k)) / (xc[i + 3] - xc[i + 1]),
                      (L(i, j + 1, k) - L(i, j - 1, k)) / (yc[j + 3] - yc[j + 1]),
                      (L(i, j, k + 1) - L(i, j, k - 1)) / (zc[k + 3] - zc[k + 1])])
        m = np.array([xc[i + 2], yc[j + 2], zc[k + 2]]) - 2 * L(i, j, k) * n
        pts = np.array([[xc[a + 2], yc[b + 2], zc[c + 2]] for a, b, c in idx])
        assert w.sum() == pytest.approx(1.0, abs=1e-12)
        # trilinear weights reproduce every linear function at the mirror point
        np.testing.assert_allclose(w @ pts, m, atol=1e-12)
        # bit order x=4, y=2, z=1 and the base cell brackets the mirror point
        assert np.array_equal(idx[7] - idx[0], [1, 1, 1]) and np.array_equal(idx[4] - idx[0], [1, 0, 0])
        assert pts[0][0] <= m[0] < pts[7][0] and pts[0][1] <= m[1] < pts[7][1]
        assert (w >= -1e-12).all()
        # mirror point lies on the fluid side, about as far out as the ghost is in
        assert np.linalg.norm(m) > 0.5


def test_flags_rule_and_invariants(I, new_oracle, cases):
    s = sphere_solver(I, new_oracle(), cases)
    lv, fl = s.get("LVSET"), s.flags()
    assert ((lv > 0) == (fl == I.FLAG_FLUID)).all()
    fluid = fl == I.FLAG_FLUID
    nb_fluid = np.zeros_like(fluid)
    for ax in range(3):
        for sh in (1, -1):
            r = np.roll(fluid, sh, axis=ax)
            sl = [slice(None)] * 3
            sl[ax] = 0 if sh == 1 else -1
            r[tuple(sl)] = False
            nb_fluid |= r
    assert (fl[(~fluid) & nb_fluid] == I.FLAG_GHOST).all()
    assert (fl[(~fluid) & ~nb_fluid] == I.FLAG_SOLID).all()
    # analytic level set: phi = |x| - 0.5 at the padded centres
    xc, yc, zc, *_ = coords(bind_oracle_hooks(s.lib) and s)
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    np.testing.assert_allclose(lv, np.sqrt(X**2 + Y**2 + Z**2) - 0.5, rtol=0, atol=1e-15)


# ---------------------------------------------------------------------------
# a5 / a8: assembled operators  <- src/ibm3d_setup.c:1566-2908
# ---------------------------------------------------------------------------
def test_operator_row_identities(I, new_oracle, cases):
    s = sphere_solver(I, new_oracle(), cases)
    bind_oracle_hooks(s.lib)
    fl = s.flags().ravel()
    P = fl.size
    ones = np.ones(P)
    y = np.zeros(P)
    inner = np.zeros(s.shape, dtype=bool)
    inner[2:-2, 2:-2, 2:-2] = True
    inner = inner.ravel()
    for t in (1, 2, 3):
        s.lib.oracle_apply_matrix(s.h, t, _p(ones), _p(y))
        np.testing.assert_allclose(y[inner & (fl == I.FLAG_FLUID)], 1.0, atol=1e-14)   # (1+sumK) - sumK
        np.testing.assert_allclose(y[inner & (fl == I.FLAG_GHOST)], 2.0, atol=1e-12)   # (1+ws) + (1-ws)
        np.testing.assert_allclose(y[inner & (fl == I.FLAG_SOLID)], 1.0, atol=0)
    s.lib.oracle_apply_matrix(s.h, 4, _p(ones), _p(y))
    np.testing.assert_allclose(y[inner & (fl != I.FLAG_SOLID)], 0.0, atol=1e-13)        # Neumann / mirror rows


def test_momentum_rows_symmetric_after_volume_scaling(I, new_oracle):
    s = new_oracle()
    bind_oracle_hooks(s.lib)
    box(I, s, [0, .3, .5, 1.1, 1.4, 2.0], [0, .2, .7, 1.0], [0, .5, .6, 1.3])
    xc, yc, zc, dx, dy, dz = coords(s)
    Ny4, Nz4 = s.Ny + 4, s.Nz + 4
    row = lambda i, j, k: ((i + 2) * Ny4 + j + 2) * Nz4 + k + 2
    cols = np.zeros(9, dtype=np.int32)
    vals = np.zeros(9)
    A = {}
    for i in range(s.Nx):
        for j in range(s.Ny):
            for k in range(s.Nz):
                n = s.lib.oracle_matrix_row(s.h, 1, row(i, j, k), cols.ctypes.data_as(_ip), _p(vals))
                A[row(i, j, k)] = dict(zip(cols[:n].tolist(), vals[:n].tolist()))
    vol = lambda i, j, k: dx[i + 2] * dy[j + 2] * dz[k + 2]
    checked = 0
    for i in range(s.Nx - 1):
        for j in range(s.Ny):
            for k in range(s.Nz):
                a, b = row(i, j, k), row(i + 1, j, k)
                assert vol(i, j, k) * A[a][b] == pytest.approx(vol(i + 1, j, k) * A[b][a], rel=1e-13)
                checked += 1
    assert checked > 20


# ---------------------------------------------------------------------------
# a2: convection  <- src/ibm3d_fracstep.c:117-170
# ---------------------------------------------------------------------------
def test_convection_exact_for_linear_field(I, new_oracle, cases):
    s = new_oracle()
    bind_oracle_hooks(s.lib)
    f = cases.uniform_faces(0.0, 1.0, 8)
    box(I, s, f, f, f, Re=100.0, dt=0.01)
    s.init_flow_const(0, 0, 0, 0)
    xc = coords(s)[0]
    a, b = 0.3, 1.7
    u1 = np.broadcast_to((a + b * xc)[:, None, None], s.shape).copy()
    s.set("U1", u1)
    s.lib.oracle_calc_N(s.h)
    N1, N2 = s.get("N1")[2:-2, 2:-2, 2:-2], s.get("N2")[2:-2, 2:-2, 2:-2]
    # N1 = d(u u)/dx = 2 u du/dx exactly for a linear u on a uniform grid
    np.testing.assert_allclose(N1, 2 * b * (a + b * xc[2:-2])[:, None, None] * np.ones_like(N1), rtol=1e-12)
    assert np.abs(N2).max() == 0.0
    # uniform flow carries no convective flux divergence
    s.set("U1", np.full(s.shape, 0.7))
    s.set("U2", np.full(s.shape, -0.2))
    s.lib.oracle_calc_N(s.h)
    assert np.abs(s.get("N1")[2:-2, 2:-2, 2:-2]).max() < 1e-14


def test_convection_conserves_momentum_in_a_periodic_box(I, new_oracle, cases):
    """calc_N is a flux difference: every face flux U u_face enters the two
    cells beside it with opposite signs, so on a periodic box the volume
    integral of N vanishes for any velocity field, here a seeded random one
    on a stretched grid. A face value interpolated differently from the two
    sides, a wrong spacing in one term or a dropped face breaks it."""
    s = new_oracle()
    bind_oracle_hooks(s.lib)
    xf = np.cumsum(np.concatenate([[0.0], 0.5 + np.linspace(0, 1, 10) ** 2]))
    yf = cases.uniform_faces(0.0, 2.0, 8)
    zf = np.cumsum(np.concatenate([[0.0], 1.0 + 0.5 * np.sin(np.arange(6))]))
    box(I, s, xf, yf, zf, Re=50.0, dt=0.01, bc=lambda s: s.set_bc(0x3f, I.BC_ALL_PERIODIC))
    s.init_flow_const(0, 0, 0, 0)
    rng = np.random.default_rng(11)
    for n in ("U1", "U2", "U3"):
        s.set(n, rng.uniform(-1, 1, s.shape))
    s.lib.oracle_update_outer(s.h)  # periodic images in the outer layers
    s.lib.oracle_calc_N(s.h)
    _, _, _, dx, dy, dz = coords(s)
    V = dx[2:-2, None, None] * dy[None, 2:-2, None] * dz[None, None, 2:-2]
    for n in ("N1", "N2", "N3"):
        N = s.get(n)[2:-2, 2:-2, 2:-2]
        scale = np.sum(np.abs(N) * V)
        assert scale > 1.0
        assert abs(np.sum(N * V)) <= 1e-13 * scale, n


def poiseuille_error(I, s, nz, steps=500, dt=0.02):
    """Channel between stationary walls at z = -1, 1 (periodic x, y), Re 1,
    driven by the body force f = 2 / Re: the steady solution is u = 1 - z^2
    (the reference's channel set-up, src/channel.c:18, 102-104). Returns the
    largest error at the cell centres after the flow has settled."""
    f4 = cases_uniform(0.0, 1.0, 4)
    zf = cases_uniform(-1.0, 1.0, nz)
    s.set_grid(f4, f4, zf)
    s.set_params(1.0, dt)
    s.set_bc(I.DIR_EAST | I.DIR_WEST | I.DIR_NORTH | I.DIR_SOUTH, I.BC_ALL_PERIODIC)
    s.set_bc(I.DIR_DOWN | I.DIR_UP, I.BC_STATIONARY_WALL)
    s.set_ext_force(2.0, 0.0, 0.0)
    s.set_obstacle(None)
    s.set_linear_solver(I.SOLVER_BiCGSTAB, I.PRECOND_AMG, 1e-10)
    s.assemble()
    s.init_flow_const(0, 0, 0, 0)
    s.iterate(steps)
    zc = (zf[1:] + zf[:-1]) / 2
    u = s.get("U1")[2:-2, 2:-2, 2:-2]
    assert np.abs(s.get("U2")[2:-2, 2:-2, 2:-2]).max() < 1e-9
    assert np.abs(s.get("U3")[2:-2, 2:-2, 2:-2]).max() < 1e-9
    return float(np.abs(u - (1 - zc ** 2)[None, None, :]).max())


def cases_uniform(a, b, n):
    return np.array([a + (b - a) * i / n for i in range(n + 1)])


def test_poiseuille_steady_state_second_order(I, new_oracle):
    """The full step (momentum RHS with the body force, the wall rows, the
    Helmholtz and pressure solves, the projection) settles to the parabolic
    profile, with the second-order error of the wall rows: the error falls by
    about 4x when the cells halve."""
    e16 = poiseuille_error(I, new_oracle(), 16)
    e32 = poiseuille_error(I, new_oracle(), 32)
    # the discrete steady state is off by h^2 / 4 (h the cell width): the
    # wall row mirrors the first cell linearly through the zero face value
    assert e16 == pytest.approx(0.125 ** 2 / 4, rel=1e-6)
    assert e32 == pytest.approx(0.0625 ** 2 / 4, rel=1e-6)


# ---------------------------------------------------------------------------
# a10: outer ghost layers  <- src/ibm3d_fracstep.c:855-1906
# ---------------------------------------------------------------------------
def test_outer_layers_linear_identities(I, new_oracle, cases):
    s = new_oracle()
    bind_oracle_hooks(s.lib)
    xf = np.array([0, .1, .25, .45, .7, 1.0, 1.4])

    def bcs(s):
        s.set_bc(I.DIR_WEST, I.BC_VELOCITY_COMPONENT, I.VAL_CONST, 0.25, -0.5, 1.0)
        s.set_bc(I.DIR_EAST, I.BC_PRESSURE, I.VAL_CONST, 0.75)
        s.set_bc(I.DIR_NORTH | I.DIR_SOUTH, I.BC_FREE_SLIP_WALL)
        s.set_bc(I.DIR_DOWN, I.BC_PRESSURE, I.VAL_CONST, 0.0)
        s.set_bc(I.DIR_UP, I.BC_STATIONARY_WALL)
    box(I, s, xf, [0, .2, .5, .9], [0, .3, .5, 1.0], bc=bcs)
    xc, yc, zc, dx, dy, dz = coords(s)
    s.init_flow_const(0, 0, 0, 0)
    lin = lambda c: (1.0 + 2.0 * c)
    for name in ("U1", "U2", "U3"):
        s.set(name, np.broadcast_to(lin(xc)[:, None, None], s.shape).copy())
    s.set("P", np.broadcast_to(lin(zc)[None, None, :], s.shape).copy())
    s.lib.oracle_update_outer(s.h)
    u1, u2, p = s.get("U1"), s.get("U2"), s.get("P")
    J, K = slice(2, -2), slice(2, -2)
    # East pressure outlet: velocity linearly extrapolated -> exact for linear u
    np.testing.assert_allclose(u1[-2, J, K], lin(xc[-2]), rtol=1e-13)
    np.testing.assert_allclose(u1[-1, J, K], lin(xc[-1]), rtol=1e-13)
    # West velocity inlet: interpolation to the face equals the BC value
    ao, bi = dx[1] / 2, dx[2] / 2
    np.testing.assert_allclose((bi * u2[1, J, K] + ao * u2[2, J, K]) / (ao + bi), -0.5, rtol=1e-13)
    ao2, bi2 = dx[0] / 2 + dx[1], dx[2] + dx[3] / 2
    np.testing.assert_allclose((bi2 * u1[0, J, K] + ao2 * u1[3, J, K]) / (ao2 + bi2), 0.25, rtol=1e-13)
    # free slip (y): tangential components copied, normal component reflected
    np.testing.assert_allclose(u1[2:-2, 1, K], u1[2:-2, 2, K])
    # Down pressure BC on p uses the x spacing (src/ibm3d_fracstep.c:1792-1800)
    a, b = dx[1] / 2, dx[2] / 2
    np.testing.assert_allclose(p[2:-2, 2:-2, 1], (a + b) / b * 0.0 - a / b * p[2:-2, 2:-2, 2], rtol=1e-13)
    # stationary wall (Up): Neumann p
    np.testing.assert_allclose(p[2:-2, 2:-2, -2], p[2:-2, 2:-2, -3])


# ---------------------------------------------------------------------------
# a1: full step on analytic flows
# ---------------------------------------------------------------------------
def taylor_green_error(I, s, n, steps, nu=0.05, dt=0.01):
    L = 2 * math.pi
    f = np.array([L * i / n for i in range(n + 1)])
    s.set_grid(f, f, f)
    s.set_params(1.0 / nu, dt)
    s.set_bc(0x3f, I.BC_ALL_PERIODIC)
    s.set_obstacle(None)
    s.set_linear_solver(I.SOLVER_BiCGSTAB, I.PRECOND_AMG, 1e-12)
    s.assemble()
    s.init_flow_const(0, 0, 0, 0)
    xc = (f[1:] + f[:-1]) / 2
    h = L / n
    xp = np.concatenate([[xc[0] - 2 * h, xc[0] - h], xc, [xc[-1] + h, xc[-1] + 2 * h]])
    X, Y, Z = np.meshgrid(xp, xp, xp, indexing="ij")
    s.set("U1", np.sin(X) * np.cos(Y))
    s.set("U2", -np.cos(X) * np.sin(Y))
    s.set("P", (np.cos(2 * X) + np.cos(2 * Y)) / 4)
    s.iterate(steps)
    t = steps * dt
    exact = np.sin(X) * np.cos(Y) * math.exp(-2 * nu * t)
    inner = (slice(2, -2),) * 3
    return np.abs(s.get("U1")[inner] - exact[inner]).max() / math.exp(-2 * nu * t)


def test_taylor_green_decay_second_order(I, new_oracle):
    e8 = taylor_green_error(I, new_oracle(), 8, 10)
    e16 = taylor_green_error(I, new_oracle(), 16, 10)
    assert e16 < 0.05
    assert 3.0 < e8 / e16 < 5.5  # spatial order 2


# ---------------------------------------------------------------------------
# a14: STL reader and closest-point transform  <- src/geo3d.c:229-704
# ---------------------------------------------------------------------------
def read_stl_triangles(path):
    d = open(path, "rb").read()
    n = int(np.frombuffer(d[80:84], dtype="<u4")[0])
    rec = np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
    r = np.frombuffer(d[84:84 + n * 50], dtype=rec)
    return r["v"].astype(np.float64), r["n"].astype(np.float64)


def point_triangle_distance(p, tri):
    """Exact distance from points p (M,3) to triangles tri (T,3,3): min over
    the interior projection, the three edges and the three vertices."""
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    n = np.cross(b - a, c - a)
    nn = np.linalg.norm(n, axis=1)
    nh = n / nn[:, None]
    d = p[:, None, :] - a[None]
    t = np.einsum("mtk,tk->mt", d, nh)
    q = p[:, None, :] - t[..., None] * nh[None]
    def inside(q):
        c1 = np.einsum("mtk,tk->mt", np.cross(b - a, q - a), n)
        c2 = np.einsum("mtk,tk->mt", np.cross(c - b, q - b), n)
        c3 = np.einsum("mtk,tk->mt", np.cross(a - c, q - c), n)
        return (c1 >= 0) & (c2 >= 0) & (c3 >= 0)
    best = np.where(inside(q), np.abs(t), np.inf)
    for u, v in ((a, b), (b, c), (c, a)):
        e = v - u
        s = np.clip(np.einsum("mtk,tk->mt", p[:, None, :] - u[None], e) / np.einsum("tk,tk->t", e, e), 0, 1)
        h = u[None] + s[..., None] * e[None]
        best = np.minimum(best, np.linalg.norm(p[:, None, :] - h, axis=2))
    return best.min(axis=1)


def test_stl_reader_counts(I, oracle_lib):
    import ctypes.util
    libc = C.CDLL(ctypes.util.find_library("c"))
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    oracle_lib.Polyhedron_read_stl.argtypes = [C.c_void_p, C.c_void_p]
    oracle_lib.oracle_poly_counts.argtypes = [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    oracle_lib.oracle_poly_counts.restype = C.c_size_t
    poly = oracle_lib.Polyhedron_new()
    fp = libc.fopen(os.path.join(GOLDEN, "sphere.stl").encode(), b"rb")
    oracle_lib.Polyhedron_read_stl(poly, fp)
    libc.fclose(fp)
    ne, nv = C.c_size_t(), C.c_size_t()
    nf = oracle_lib.oracle_poly_counts(poly, C.byref(ne), C.byref(nv))
    assert (nf, ne.value, nv.value) == (528, 792, 266)   # closed: V - E + F = 2
    oracle_lib.Polyhedron_destroy(poly)


def test_cpt_matches_brute_force_distance(I, new_oracle, cases, oracle_lib):
    import ctypes.util
    libc = C.CDLL(ctypes.util.find_library("c"))
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    oracle_lib.Polyhedron_read_stl.argtypes = [C.c_void_p, C.c_void_p]
    poly = oracle_lib.Polyhedron_new()
    fp = libc.fopen(os.path.join(GOLDEN, "sphere.stl").encode(), b"rb")
    oracle_lib.Polyhedron_read_stl(poly, fp)
    libc.fclose(fp)
    s = new_oracle()
    f = cases.uniform_faces(-0.9, 0.9, 12)
    s.set_grid(f, f, f)
    s.set_params(100, 0.01)
    s.set_bc(0x3f, I.BC_STATIONARY_WALL)
    s.set_obstacle(poly)
    s.set_linear_solver(I.SOLVER_BiCGSTAB, I.PRECOND_AMG, 1e-8)
    s.assemble()
    oracle_lib.Polyhedron_destroy(poly)
    lv = s.get("LVSET")
    xc, yc, zc, *_ = coords(bind_oracle_hooks(s.lib) and s)
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    tri, nrm = read_stl_triangles(os.path.join(GOLDEN, "sphere.stl"))
    dist = point_triangle_distance(P, tri)
    # convex closed mesh: a point is inside iff it is behind every face plane
    side = np.einsum("mtk,tk->mt", P[:, None, :] - tri[None, :, 0], nrm)
    inside = (side <= 0).all(axis=1)
    sd = np.where(inside, -dist, dist)
    np.testing.assert_allclose(lv.ravel(), sd, atol=1e-6)
    assert ((lv.ravel() > 0) == (sd > 0)).all()


# ---------------------------------------------------------------------------
# regression fixture written by tests/golden/make_golden.py
# ---------------------------------------------------------------------------
def test_cavity_golden_fixture(I, new_oracle, cases):
    path = os.path.join(GOLDEN, "cavity8_5steps.npz")
    g = np.load(path)
    s = new_oracle()
    cases.setup_cavity(s, n=8, tol=1e-12)
    s.iterate(5)
    for name in ("U1", "U2", "U3", "P"):
        np.testing.assert_allclose(s.get(name), g[name], rtol=0, atol=1e-10)


# ---------------------------------------------------------------------------
# a3: Smagorinsky (src/ibm3d_turb.c:8-45) and the van Driest extension
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("model,cs", [(1, 0.1), (1, 1.0), (4, 0.1), (4, 1.0)])
def test_smagorinsky_closed_form_linear_shear(I, new_oracle, cases, model, cs):
    """u = a y: central differences are exact, S_12 = a/2, |S| = |a|, so
    tau_12 = -2 l^2 |a| (a/2) with l = min(Cs Delta f, 0.41 phi), Delta =
    2 (dx dy dz)^(1/3), phi = 0.5 without a body, f = 1 (Smagorinsky) or
    1 - exp(-y+/A+) with y+ = (distance to the y walls) u_tau Re."""
    a, Re, aplus, utau = 3.0, 50.0, 26.0, 0.8
    xf = cases.uniform_faces(0.0, 1.0, 6)
    yf = np.array([-math.cos(math.pi * j / 8) for j in range(9)])
    zf = cases.uniform_faces(0.0, 0.5, 4)
    s = new_oracle()

    def bc(s):
        s.set_bc(I.DIR_EAST | I.DIR_WEST | I.DIR_NORTH | I.DIR_SOUTH | I.DIR_DOWN | I.DIR_UP,
                 I.BC_STATIONARY_WALL)
        s.set_turb_model(model, cs, aplus=aplus, u_tau=utau, walls=I.DIR_NORTH | I.DIR_SOUTH)

    box(I, s, xf, yf, zf, Re=Re, dt=0.01, bc=bc)
    bind_oracle_hooks(s.lib)
    xc, yc, zc, dx, dy, dz = coords(s)
    s.init_flow_const(0, 0, 0, 0)
    s.set("U1", np.broadcast_to(a * yc[None, :, None], s.shape).copy())
    s.lib.IBMSolver_calc_tau_r(s.h)
    t12 = s.get("TAU12")[2:-2, 2:-2, 2:-2]
    D = 2 * np.cbrt(dx[2:-2, None, None] * dy[None, 2:-2, None] * dz[None, None, 2:-2])
    l = cs * D
    if model == 4:
        yw = np.minimum(yc[2:-2] - yf[0], yf[-1] - yc[2:-2])
        l = l * (1 - np.exp(-(yw * utau * Re) / aplus))[None, :, None]
    l = np.minimum(l, 0.41 * 0.5)
    expect = -2 * l**2 * abs(a) * (a / 2)
    np.testing.assert_allclose(t12, np.broadcast_to(expect, t12.shape), rtol=1e-12, atol=0)
    assert np.allclose(s.get("TAU13")[2:-2, 2:-2, 2:-2], 0) and np.allclose(s.get("TAU11")[2:-2, 2:-2, 2:-2], 0)


def test_vreman_closed_form_constant_gradient(I, new_oracle, cases):
    """Vreman (extension; Vreman, Phys. Fluids 16, 3670, 2004) on a velocity
    field with a constant, non-symmetric gradient alpha_mn = du_n/dx_m on a
    grid stretched in y: central differences are exact for linear fields, so
    per cell nu_t = c sqrt(B / |alpha|^2) with beta = alpha^T D^2 alpha (D =
    diag of the cell widths) and B = ((tr beta)^2 - tr(beta^2)) / 2, the second
    invariant of beta (the oracle sums the 2x2 principal minors instead);
    tau = -2 nu_t S. A transposed alpha, a missing width or the cube-root
    filter width of Smagorinsky all change tau here."""
    c = 0.07
    G = np.array([[0.3, -1.1, 0.4], [2.0, 0.2, -0.7], [0.5, 0.9, -0.5]])  # G[m, n] = du_n/dx_m
    xf = cases.uniform_faces(0.0, 1.2, 6)
    yf = np.array([-math.cos(math.pi * j / 8) for j in range(9)])
    zf = cases.uniform_faces(0.0, 0.5, 4)
    s = new_oracle()

    def bc(s):
        s.set_bc(I.DIR_EAST | I.DIR_WEST | I.DIR_NORTH | I.DIR_SOUTH | I.DIR_DOWN | I.DIR_UP,
                 I.BC_STATIONARY_WALL)
        s.set_turb_model(I.TURBMODEL_VREMAN, c)

    box(I, s, xf, yf, zf, Re=100.0, dt=0.01, bc=bc)
    bind_oracle_hooks(s.lib)
    xc, yc, zc, dx, dy, dz = coords(s)
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    s.init_flow_const(0, 0, 0, 0)
    for n, name in enumerate(("U1", "U2", "U3")):
        s.set(name, G[0, n] * X + G[1, n] * Y + G[2, n] * Z)
    s.lib.IBMSolver_calc_tau_r(s.h)
    S = 0.5 * (G + G.T)
    got = {nm: s.get(nm)[2:-2, 2:-2, 2:-2] for nm in ("TAU11", "TAU12", "TAU13", "TAU22", "TAU23", "TAU33")}
    for i in range(len(xf) - 1):
        for j in range(len(yf) - 1):
            for k in range(len(zf) - 1):
                D2 = np.diag([dx[i + 2] ** 2, dy[j + 2] ** 2, dz[k + 2] ** 2])
                beta = G.T @ D2 @ G
                B = 0.5 * (np.trace(beta) ** 2 - np.trace(beta @ beta))
                nut = c * math.sqrt(B / np.sum(G * G))
                for (a, b), nm in zip(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)), got):
                    assert got[nm][i, j, k] == pytest.approx(-2 * nut * S[a, b], rel=1e-12, abs=0)
    # hand-computed cell: alpha = diag(a1, a2, a3) gives B = sum over pairs of
    # (D_m a_m D_n a_n)^2
    s.set("U1", 1.5 * X)
    s.set("U2", -0.5 * Y)
    s.set("U3", -1.0 * Z)
    s.lib.IBMSolver_calc_tau_r(s.h)
    i, j, k = 1, 0, 2
    d = [dx[i + 2] * 1.5, dy[j + 2] * 0.5, dz[k + 2] * 1.0]
    B = (d[0] * d[1]) ** 2 + (d[0] * d[2]) ** 2 + (d[1] * d[2]) ** 2
    nut = c * math.sqrt(B / (1.5 ** 2 + 0.5 ** 2 + 1.0))
    assert s.get("TAU11")[i + 2, j + 2, k + 2] == pytest.approx(-2 * nut * 1.5, rel=1e-13)
    assert s.get("TAU22")[i + 2, j + 2, k + 2] == pytest.approx(-2 * nut * -0.5, rel=1e-13)
    # pure shear u = a y (one non-zero alpha): B = 0, so Vreman switches off
    s.set("U1", 3.0 * Y)
    s.set("U2", 0 * Y)
    s.set("U3", 0 * Y)
    s.lib.IBMSolver_calc_tau_r(s.h)
    for nm in got:
        assert np.all(s.get(nm)[2:-2, 2:-2, 2:-2] == 0), nm


# ---------------------------------------------------------------------------
# a11: in-place ghost-cell update  <- src/ibm3d_fracstep.c:1915-1976
# ---------------------------------------------------------------------------
def plane_body(I, s, cases, normal, a, n=(12, 10, 9)):
    """A box cut by the plane normal . x = a (|normal| = 1, the body on the
    side normal . x < a), level set installed directly: phi = normal . x - a
    is the signed distance, so the ghost stencils' mirror points are exact."""
    xf = cases.uniform_faces(0.0, 1.2, n[0])
    yf = 0.5 * (1 - np.cos(np.pi * np.arange(n[1] + 1) / n[1]))
    zf = cases.uniform_faces(0.0, 0.9, n[2])
    box(I, s, xf, yf, zf, Re=10.0, dt=0.01)
    lib = bind_oracle_hooks(s.lib)
    lib.oracle_set_lvset.argtypes = [C.c_void_p, _dp]
    xc, yc, zc, *_ = coords(s)
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    nv = np.asarray(normal, dtype=float) / np.linalg.norm(normal)
    phi = np.ascontiguousarray(nv[0] * X + nv[1] * Y + nv[2] * Z - a)
    lib.oracle_set_lvset(s.h, _p(phi))
    return X, Y, Z, phi, nv


PLANES = [((1.0, 0.3, 0.2), 0.45), ((-1.0, 2.0, -2.0), -0.6), ((-3.0, -1.0, 0.5), -1.0)]


@pytest.mark.parametrize("normal,a", PLANES)
def test_update_ghost_exact_about_a_tilted_plane(I, new_oracle, cases, normal, a):
    """Mirror-image boundary conditions are exact for linear fields about a
    plane: a field odd about it (u = c phi) gives u_g = -u(mirror) = c phi_g,
    a field even about it (p linear along the plane) gives p_g = p(mirror) =
    p(x_g). The exact ghost values are a fixed point of the pass, and from
    zero ghost values repeated passes converge to them (the self weight on the
    left-hand side with the sign of each condition: 1 + w for u, 1 - w for
    p). Solid cells become NaN."""
    s = new_oracle()
    X, Y, Z, phi, nv = plane_body(I, s, cases, normal, a)
    fl = s.flags()
    ghost, solid = fl == I.FLAG_GHOST, fl == I.FLAG_SOLID
    assert ghost[2:-2, 2:-2, 2:-2].sum() > 40 and solid[2:-2, 2:-2, 2:-2].sum() > 40
    t = np.cross(nv, [0.3, -0.2, 1.0])  # tangential direction
    exact = {"U1": 2.0 * phi, "U2": -0.7 * phi, "U3": 1.3 * phi,
             "P": 0.25 + t[0] * X + t[1] * Y + t[2] * Z}
    for nm, v in exact.items():
        s.set(nm, v)
    s.lib.oracle_update_ghost(s.h)
    inner = np.zeros(fl.shape, bool)
    inner[2:-2, 2:-2, 2:-2] = True
    for nm, v in exact.items():
        got = s.get(nm)
        np.testing.assert_allclose(got[ghost & inner], v[ghost & inner], rtol=0, atol=1e-13)
        assert np.isnan(got[solid & inner]).all()
    # from zero ghost values
    for nm, v in exact.items():
        start = v.copy()
        start[ghost & inner] = 0.0
        s.set(nm, start)
    for _ in range(40):
        s.lib.oracle_update_ghost(s.h)
    for nm, v in exact.items():
        np.testing.assert_allclose(s.get(nm)[ghost & inner], v[ghost & inner], rtol=0, atol=1e-11)


@pytest.mark.parametrize("normal,a", PLANES)
def test_update_ghost_in_place_lexicographic_order(I, new_oracle, cases, normal, a):
    """One pass from arbitrary ghost values equals a serial sweep over the
    interior in (i, j, k) order, k fastest, in which a ghost stencil reads the
    values already written to ghost cells earlier in that order and the old
    values of later ones; solid stencil points are dropped and the weights
    renormalised (src/ibm3d_fracstep.c:1915-1976). Brute force on the CPU with
    the exported stencil routine (pinned by the mirror-point test). A
    simultaneous (Jacobi) update differs on these planes, so the check sees
    the order."""
    s = new_oracle()
    plane_body(I, s, cases, normal, a)
    fl = s.flags()
    rng = np.random.default_rng(11)
    vals = {nm: rng.uniform(-1, 1, s.shape) for nm in ("U1", "U2", "U3", "P")}
    for nm, v in vals.items():
        s.set(nm, v)
    s.lib.oracle_update_ghost(s.h)
    got = {nm: s.get(nm) for nm in vals}

    idx = np.zeros((8, 3), dtype=np.int32)
    w = np.zeros(8)
    seq = {nm: v.copy() for nm, v in vals.items()}
    jac = {nm: v.copy() for nm, v in vals.items()}
    for (i, j, k) in s.ghost_cells():  # lexicographic, k fastest
        s.lib.IBMSolver_ghost_interp(s.h, int(i), int(j), int(k), idx.ctypes.data_as(_ip), _p(w))
        keep = np.array([fl[a + 2, b + 2, c + 2] != I.FLAG_SOLID for a, b, c in idx])
        wn = np.where(keep, w, 0.0) / w[keep].sum()
        me = np.array([(a, b, c) == (i, j, k) for a, b, c in idx])
        ws = wn[me].sum()
        others = [(wn[l], tuple(idx[l] + 2)) for l in range(8) if keep[l] and not me[l]]
        for out, src in ((seq, seq), (jac, vals)):
            for nm in vals:
                acc = sum(wl * src[nm][q] for wl, q in others)
                if nm == "P":
                    if 1 - ws != 0:
                        out[nm][i + 2, j + 2, k + 2] = acc / (1 - ws)
                else:
                    out[nm][i + 2, j + 2, k + 2] = -acc / (1 + ws)
    g = fl == I.FLAG_GHOST
    differs = False
    for nm in vals:
        np.testing.assert_allclose(got[nm][g], seq[nm][g], rtol=1e-13, atol=1e-14)
        differs |= not np.allclose(got[nm][g], jac[nm][g], rtol=1e-10, atol=1e-12)
    # where the body lies on the high side of an axis, mirror stencils reach
    # ghost cells that come earlier in the sweep
    assert differs == (min(normal) < 0)


# ---------------------------------------------------------------------------
# drag / lift: force on the body from the staircase ghost-cell surface
# (extension, no reference counterpart)
# ---------------------------------------------------------------------------
def body_force_closed_forms(I, s, cases, check):
    """Closed forms of the staircase surface integral. The faces between
    interior fluid and ghost cells bound the union B of the non-fluid cells,
    so by the divergence theorem on B (a union of boxes):
      p = g . x (any grid; the face pressure is interpolated to the face):
        F = -g |B|;
      u_b = S x_a^2 / 2 on a uniform grid (the centre difference across a face
        is the derivative at the face): F = (1/Re) S |B| e_b;
      u_b = S x_a: F = 0.
    `s` is a fresh solver; `check(got, expect)` compares."""
    Re = 50.0
    # stretched grid: pressure part
    xf, yf, zf = cases.sphere_faces(24, 20, 20, 0.12)
    cases.setup_sphere(s[0], faces=(xf, yf, zf), Re=Re)
    fl = s[0].flags()[2:-2, 2:-2, 2:-2]
    d = [np.diff(f) for f in (xf, yf, zf)]
    vol = (d[0][:, None, None] * d[1][None, :, None] * d[2][None, None, :])[fl != I.FLAG_FLUID].sum()
    assert (fl[[0, -1], :, :] == I.FLAG_FLUID).all() and (fl[:, [0, -1], :] == I.FLAG_FLUID).all()
    xc, yc, zc = (cases.cell_centres(f) for f in (xf, yf, zf))
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    g = np.array([0.7, -1.3, 0.4])
    for nm in ("U1", "U2", "U3"):
        s[0].set(nm, 0 * X)
    s[0].set("P", 2.0 + g[0] * X + g[1] * Y + g[2] * Z)
    check(s[0].body_force(), -g * vol)
    # uniform grid: viscous part
    f = cases.uniform_faces(-1.5, 1.5, 24)
    cases.setup_sphere(s[1], faces=(f, f, f), Re=Re)
    fl = s[1].flags()[2:-2, 2:-2, 2:-2]
    vol = (fl != I.FLAG_FLUID).sum() * (3.0 / 24) ** 3
    c = cases.cell_centres(f)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    S = 1.7
    s[1].set("P", 0 * X)
    s[1].set("U1", 0 * X)
    s[1].set("U3", 0 * X)
    s[1].set("U2", S * X**2 / 2)  # u_y = S x^2 / 2: F = S |B| / Re e_y
    check(s[1].body_force(), np.array([0.0, S * vol / Re, 0.0]))
    s[1].set("U2", 0 * X)
    s[1].set("U3", S * Y)  # uniform shear: no net force
    check(s[1].body_force(), np.zeros(3))


def test_body_force_closed_forms(I, new_oracle, cases):
    def check(got, expect):
        np.testing.assert_allclose(got, expect, rtol=1e-12, atol=1e-13)
    body_force_closed_forms(I, [new_oracle(), new_oracle()], cases, check)
