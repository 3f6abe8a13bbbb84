"""Workload definitions: grids, boundary conditions and initial fields.

Each builder configures an ``ibm3d.Solver`` through the reference's C API
exactly like the reference drivers do (src/main.c:133-159, src/cavity.c:50-
77): grid, params, BCs, obstacle, linear solver, assemble, init.

Workloads:
  cavity   3-D lid-driven cavity on the unit cube, lid at y = 1 (DIR_NORTH)
           moving with u = 1, stationary walls elsewhere; u = v = w = p = 0
           plus a seeded uniform perturbation of amplitude 1e-3 inside.
  sphere   flow past a sphere of diameter 1 at the origin, phi = |x| - 0.5,
           domain [-5,15]x[-5,5]^2, cosine-stretched toward the body; West
           velocity inlet (1,0,0), East pressure outlet p = 0, free-slip
           lateral walls; initial u = (1,0,0), p = 0 (src/main.c:139-155).
"""
from __future__ import annotations

import math
import struct

import numpy as np

from . import ibm3d as I

MASK = (1 << 64) - 1


# ---------------------------------------------------------------------------
# seeded perturbation: counter-based SplitMix64 over the bit patterns of the
# cell-centre coordinates, so the draw is identical on every backend and does
# not depend on traversal order.
# ---------------------------------------------------------------------------
def _sm64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & MASK
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def _bits(d: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", d))[0]


def uniform01(seed: int, comp: int, x: float, y: float, z: float) -> float:
    h = _sm64(_bits(z))
    h = _sm64(_bits(y) ^ h)
    h = _sm64(_bits(x) ^ h)
    h = _sm64(comp ^ h)
    h = _sm64(seed ^ h)
    return (h >> 11) * (1.0 / (1 << 53))


def _sm64_np(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01_grid(seed: int, comp: int, xc, yc, zc) -> np.ndarray:
    """Vectorised uniform01 over the tensor grid xc x yc x zc (k fastest)."""
    bx = np.asarray(xc, dtype=np.float64).view(np.uint64)[:, None, None]
    by = np.asarray(yc, dtype=np.float64).view(np.uint64)[None, :, None]
    bz = np.asarray(zc, dtype=np.float64).view(np.uint64)[None, None, :]
    h = _sm64_np(np.broadcast_to(bz, (len(xc), len(yc), len(zc))).copy())
    h = _sm64_np(by ^ h)
    h = _sm64_np(bx ^ h)
    h = _sm64_np(np.uint64(comp) ^ h)
    h = _sm64_np(np.uint64(seed) ^ h)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


# ---------------------------------------------------------------------------
# grids
# ---------------------------------------------------------------------------
def uniform_faces(a: float, b: float, n: int) -> np.ndarray:
    return np.array([a + (b - a) * i / n for i in range(n + 1)])


def cosspace(a: float, b: float, n: int) -> np.ndarray:
    """Cosine clustering toward both ends (the reference's cosspace.c:14-17)."""
    return np.array([(b - a) * (1 - math.cos(math.pi * i / n)) / 2 + a for i in range(n + 1)])


def _partial_cos(L: float, n: int, theta0: float) -> np.ndarray:
    """n cells on [0, L] clustered toward 0 by a cosine over [theta0, pi/2]:
    x(s) = L (cos t0 - cos(t0 + s (pi/2 - t0))) / cos t0."""
    s = np.arange(n + 1) / n
    th = theta0 + s * (math.pi / 2 - theta0)
    x = L * (math.cos(theta0) - np.cos(th)) / math.cos(theta0)
    x[0], x[-1] = 0.0, L
    return x


def toward_point_faces(a: float, b: float, n: int, c: float, dmin: float) -> np.ndarray:
    """Faces on [a, b] clustered toward the interior point c with smallest cell
    width ~dmin on both sides: a partial-angle cosine map per side, the cell
    count split so that both sides start from the same width. This is the
    pinned stretching of the sphere workload (the reference's cosspace.c only
    clusters toward the domain ends)."""
    La, Lb = c - a, b - c

    def theta_for(L, m):
        lo, hi = 1e-6, math.pi / 2 - 1e-6
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if _partial_cos(L, m, mid)[1] < dmin:
                lo = mid
            else:
                hi = mid
        return 0.5 * (lo + hi)

    best = None
    for na in range(2, n - 1):
        nb = n - na
        ta, tb = theta_for(La, na), theta_for(Lb, nb)
        xa, xb = _partial_cos(La, na, ta), _partial_cos(Lb, nb, tb)
        # match the largest widths of the two sides as well as possible
        score = abs(math.log((xa[-1] - xa[-2]) / (xb[-1] - xb[-2])))
        if best is None or score < best[0]:
            best = (score, xa, xb)
    _, xa, xb = best
    faces = np.concatenate([c - xa[::-1], c + xb[1:]])
    faces[0], faces[-1] = a, b
    return faces


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
CAVITY_SEED = 2020


def setup_cavity(s: I.Solver, n: int = 32, Re: float = 100.0, dt: float = 0.005,
                 tol: float = 1e-10, seed: int = CAVITY_SEED, amp: float = 1e-3,
                 precision: int = I.PREC_FP64, init: str = "array", shape=None,
                 solver: int = I.SOLVER_BiCGSTAB, precond: int = I.PRECOND_AMG, assemble: bool = True):
    """Lid-driven cavity on the unit cube, n^3 cells (or shape = (nx, ny, nz)).
    assemble=False stops before IBMSolver_assemble and returns the initial
    flow as an IBMServeInit (rest plus the seeded perturbation of the
    interior cells, the cells whose centres lie inside the cube), for a
    serving job that assembles and initialises natively."""
    nx, ny, nz = shape if shape is not None else (n, n, n)
    fx, fy, fz = uniform_faces(0.0, 1.0, nx), uniform_faces(0.0, 1.0, ny), uniform_faces(0.0, 1.0, nz)
    s.set_grid(fx, fy, fz)
    s.set_params(Re, dt)
    s.set_bc(I.DIR_NORTH, I.BC_VELOCITY_COMPONENT, I.VAL_CONST, 1.0, 0.0, 0.0)
    s.set_bc(I.DIR_EAST | I.DIR_WEST | I.DIR_SOUTH | I.DIR_DOWN | I.DIR_UP, I.BC_STATIONARY_WALL)
    s.set_obstacle(None)
    s.set_linear_solver(solver, precond, tol)
    s.set_precision(precision)
    if not assemble:
        return I.ServeInit(0.0, 0.0, 0.0, 0.0, amp, seed)
    s.assemble()

    def pert(comp):
        def f3(x, y, z):
            if 0.0 < x < 1.0 and 0.0 < y < 1.0 and 0.0 < z < 1.0:
                return amp * (2.0 * uniform01(seed, comp, x, y, z) - 1.0)
            return 0.0
        return f3

    if init == "func":
        s.init_flow_func(pert(1), pert(2), pert(3), lambda x, y, z: 0.0)
    else:
        s.init_flow_const(0.0, 0.0, 0.0, 0.0)
        xc, yc, zc = cell_centres(fx), cell_centres(fy), cell_centres(fz)
        inside = (((xc > 0) & (xc < 1))[:, None, None] & ((yc > 0) & (yc < 1))[None, :, None]
                  & ((zc > 0) & (zc < 1))[None, None, :])
        for comp, name in ((1, "U1"), (2, "U2"), (3, "U3")):
            a = amp * (2.0 * uniform01_grid(seed, comp, xc, yc, zc) - 1.0)
            s.set(name, np.where(inside, a, 0.0))
    return s


def cell_centres(faces: np.ndarray, periodic: bool = False) -> np.ndarray:
    """Padded cell centres (2 ghost layers) as IBMSolver_assemble builds them
    for non-periodic axes (src/ibm3d_setup.c:536-588)."""
    f = np.asarray(faces, dtype=np.float64)
    n = len(f) - 1
    d = np.zeros(n + 4)
    c = np.zeros(n + 4)
    d[2:n + 2] = f[1:] - f[:-1]
    c[2:n + 2] = (f[1:] + f[:-1]) / 2
    if periodic:
        d[1], d[0], d[n + 2], d[n + 3] = d[n + 1], d[n], d[2], d[3]
    else:
        d[1], d[0], d[n + 2], d[n + 3] = d[2], d[3], d[n + 1], d[n]
    c[1] = c[2] - (d[1] + d[2]) / 2
    c[0] = c[1] - (d[0] + d[1]) / 2
    c[n + 2] = c[n + 1] + (d[n + 1] + d[n + 2]) / 2
    c[n + 3] = c[n + 2] + (d[n + 2] + d[n + 3]) / 2
    return c


SPHERE_DMIN = 0.015


def sphere_faces(nx: int = 256, ny: int = 128, nz: int = 128, dmin: float = SPHERE_DMIN):
    return (toward_point_faces(-5.0, 15.0, nx, 0.0, dmin),
            toward_point_faces(-5.0, 5.0, ny, 0.0, dmin),
            toward_point_faces(-5.0, 5.0, nz, 0.0, dmin))


def setup_sphere(s: I.Solver, nx: int = 256, ny: int = 128, nz: int = 128, Re: float = 300.0,
                 dt: float = 0.002, tol: float = 1e-10, precision: int = I.PREC_FP64,
                 dmin: float | None = None, faces=None, solver: int = I.SOLVER_BiCGSTAB,
                 precond: int = I.PRECOND_AMG, turb: tuple | None = None, outflow: str = "pressure",
                 assemble: bool = True):
    """Flow past the analytic sphere phi = |x| - 0.5: velocity inlet (1, 0, 0)
    on the West face, free-slip lateral walls, and on the East face a
    pressure outlet p = 0 (outflow "pressure", the reference's BC_PRESSURE) or
    a convective outflow advected with the mean outlet velocity ("convective",
    BC_CONVECTIVE_OUTFLOW); turb = (model, coefficient) switches on an
    eddy-viscosity model. assemble=False stops before IBMSolver_assemble and
    returns the initial flow (uniform u = 1) as an IBMServeInit."""
    if faces is None:
        if dmin is None:
            # keep the cell-width ratio of the 256x128x128 grid on coarser grids
            dmin = SPHERE_DMIN * 128.0 / ny
        faces = sphere_faces(nx, ny, nz, dmin)
    xf, yf, zf = faces
    s.set_grid(xf, yf, zf)
    s.set_params(Re, dt)
    s.set_bc(I.DIR_WEST, I.BC_VELOCITY_COMPONENT, I.VAL_CONST, 1.0, 0.0, 0.0)
    if outflow == "convective":
        s.set_bc(I.DIR_EAST, I.BC_CONVECTIVE_OUTFLOW, 0.0)
    elif outflow == "pressure":
        s.set_bc(I.DIR_EAST, I.BC_PRESSURE, I.VAL_CONST, 0.0)
    else:
        raise ValueError(f"unknown outflow {outflow!r}")
    s.set_bc(I.DIR_NORTH | I.DIR_SOUTH | I.DIR_DOWN | I.DIR_UP, I.BC_FREE_SLIP_WALL)
    s.set_obstacle_sphere(0.0, 0.0, 0.0, 0.5)
    if turb is not None:
        s.set_turb_model(*turb)
    s.set_linear_solver(solver, precond, tol)
    s.set_precision(precision)
    if not assemble:
        return I.ServeInit(1.0, 0.0, 0.0, 0.0, 0.0, 0)
    s.assemble()
    s.init_flow_const(1.0, 0.0, 0.0, 0.0)
    return s


# ---------------------------------------------------------------------------
# LES channel at Re_tau = 180 in the reference's units (src/channel.c:13-18,
# 47-60, 102-104, with y as the wall-normal axis): periodic x and z,
# stationary walls at y = +-1, y_j = -cos(pi j / Ny), velocity scaled by the
# laminar centreline velocity (u = 1 - y^2), Re = 3300 (the classic
# centreline Reynolds number of the Re_tau = 180 channel), constant mean
# pressure gradient f = 2 / Re (the one that holds the laminar profile).
# ---------------------------------------------------------------------------
CHANNEL_SEED = 3
CHANNEL_RE = 3300.0
CHANNEL_RETAU = 180.0
CHANNEL_BULK = 2.0 / 3.0  # bulk velocity of the laminar parabola


def channel_faces(nx: int = 192, ny: int = 96, nz: int = 128):
    return (uniform_faces(0.0, 4 * math.pi, nx),
            np.array([-math.cos(math.pi * j / ny) for j in range(ny + 1)]),
            uniform_faces(0.0, 4 * math.pi / 3, nz))


def channel_initial(xc, yc, zc, seed: int = CHANNEL_SEED, bulk: float = CHANNEL_BULK, amp: float = 0.1,
                    smooth: int = 4):
    """Laminar parabola u = 1.5 bulk (1 - y^2) plus seeded Gaussian noise
    made solenoidal: the noise is the curl of a vector potential whose
    components are independent Gaussian draws (one per padded cell, seed),
    smoothed by `smooth` passes of the [1, 2, 1] / 4 filter along each axis
    (periodic in x and z; so that the noise is resolved by the grid rather
    than a cell-to-cell oscillation) and damped by (1 - y^2)^2 toward the
    walls; the curl by central differences (periodic in x and z), scaled to
    an RMS of amp * bulk over the interior. Arrays are over the padded cell
    centres (the reference layout)."""
    rng = np.random.default_rng(seed)
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    damp = np.clip(1 - Y**2, 0.0, None) ** 2

    def filt(f):
        for _ in range(smooth):
            for axis in (0, 2):
                f = 0.25 * np.roll(f, 1, axis) + 0.5 * f + 0.25 * np.roll(f, -1, axis)
            g = f.copy()
            g[:, 1:-1, :] = 0.25 * f[:, :-2, :] + 0.5 * f[:, 1:-1, :] + 0.25 * f[:, 2:, :]
            f = g
        return f
    psi = [damp * filt(rng.standard_normal(X.shape)) for _ in range(3)]

    def d(f, axis, c):
        # central difference on the padded array, periodic wrap where it runs
        # past the ends (the padding layers are rewritten by the solver)
        fp, fm = np.roll(f, -1, axis), np.roll(f, 1, axis)
        cc = np.asarray(c)
        h = np.roll(cc, -1) - np.roll(cc, 1)
        h[0], h[-1] = cc[1] - cc[0], cc[-1] - cc[-2]
        shape = [1, 1, 1]
        shape[axis] = -1
        return (fp - fm) / h.reshape(shape)
    u = d(psi[2], 1, yc) - d(psi[1], 2, zc)
    v = d(psi[0], 2, zc) - d(psi[2], 0, xc)
    w = d(psi[1], 0, xc) - d(psi[0], 1, yc)
    inner = (slice(2, -2),) * 3
    rms = math.sqrt(np.mean(u[inner] ** 2 + v[inner] ** 2 + w[inner] ** 2) / 3)
    s = amp * bulk / rms
    return 1.5 * bulk * np.clip(1 - Y**2, 0.0, None) + s * u, s * v, s * w


def setup_channel(s: I.Solver, nx: int = 192, ny: int = 96, nz: int = 128, dt: float = 0.002,
                  tol: float = 1e-9, precision: int = I.PREC_FP64, turb: tuple | None = None,
                  seed: int = CHANNEL_SEED):
    """The LES channel workload: turb = (TURBMODEL_SMAGORINSKY_VD, Cs) for
    Smagorinsky with van Driest damping at the y walls (A+ = 26, u_tau =
    Re_tau / Re, so that y+ = Re_tau y_wall), (TURBMODEL_VREMAN, c) for
    Vreman, (TURBMODEL_SMAGORINSKY, Cs) for the reference's model, None for
    no model."""
    xf, yf, zf = channel_faces(nx, ny, nz)
    s.set_grid(xf, yf, zf)
    s.set_params(CHANNEL_RE, dt)
    s.set_bc(I.DIR_WEST | I.DIR_EAST | I.DIR_DOWN | I.DIR_UP, I.BC_ALL_PERIODIC)
    s.set_bc(I.DIR_NORTH | I.DIR_SOUTH, I.BC_STATIONARY_WALL)
    s.set_ext_force(2.0 / CHANNEL_RE, 0.0, 0.0)
    if turb is not None:
        model, coef = turb
        if model == I.TURBMODEL_SMAGORINSKY_VD:
            s.set_turb_model(model, coef, aplus=26.0, u_tau=CHANNEL_RETAU / CHANNEL_RE,
                             walls=I.DIR_NORTH | I.DIR_SOUTH)
        else:
            s.set_turb_model(model, coef)
    s.set_obstacle(None)
    s.set_linear_solver(I.SOLVER_BiCGSTAB, I.PRECOND_AMG, tol)
    s.set_precision(precision)
    s.assemble()
    xc = cell_centres(xf, periodic=True)
    yc = cell_centres(yf)
    zc = cell_centres(zf, periodic=True)
    u, v, w = channel_initial(xc, yc, zc, seed)
    s.init_flow_const(0, 0, 0, 0)
    s.set("U1", u)
    s.set("U2", v)
    s.set("U3", w)
    return s


# ---------------------------------------------------------------------------
# Urban canopy: 16 axis-aligned blocks on the ground of a 16 x 8 x 4 domain
# (x streamwise, y spanwise, z up), triangulated to a binary STL in memory and
# read through Polyhedron_read_stl as the reference's building driver reads
# its STL (src/building.c:61-75), log-law inflow by VAL_FUNC, Vreman LES.
# ---------------------------------------------------------------------------
URBAN_SEED = 4
URBAN_DOMAIN = (16.0, 8.0, 4.0)
URBAN_USTAR, URBAN_Z0, KARMAN = 0.05, 0.01, 0.41
URBAN_BASE = -0.5  # blocks reach below the ground so no ghost stencil crosses it


def urban_blocks(seed: int = URBAN_SEED, n: int = 16, gap: float = 0.3, x_range=(1.5, 12.5)):
    """n blocks (x0, x1, y0, y1, height): footprints U[0.5, 1.5]^2, heights
    U[0.5, 2.0], centres uniform over x_range x [0, 8] with the footprint
    inside the domain, at least `gap` apart (rejection sampling, seeded)."""
    rng = np.random.default_rng(seed)
    Ly = URBAN_DOMAIN[1]
    out = []
    while len(out) < n:
        wx, wy = rng.uniform(0.5, 1.5, 2)
        h = rng.uniform(0.5, 2.0)
        cx = rng.uniform(x_range[0] + wx / 2, x_range[1] - wx / 2)
        cy = rng.uniform(wy / 2 + gap, Ly - wy / 2 - gap)
        b = (cx - wx / 2, cx + wx / 2, cy - wy / 2, cy + wy / 2, h)
        if all(b[0] >= o[1] + gap or o[0] >= b[1] + gap or b[2] >= o[3] + gap or o[2] >= b[3] + gap for o in out):
            out.append(b)
    return out


def blocks_stl(blocks, base: float = URBAN_BASE) -> bytes:
    """Binary STL of closed boxes: 12 triangles each, counter-clockwise seen
    from outside, normals outward."""
    tris = []
    for (x0, x1, y0, y1, h) in blocks:
        z0, z1 = base, h
        v = [(x, y, z) for z in (z0, z1) for y in (y0, y1) for x in (x0, x1)]  # index x + 2y + 4z
        quads = [((0, 2, 3, 1), (0, 0, -1)), ((4, 5, 7, 6), (0, 0, 1)), ((0, 1, 5, 4), (0, -1, 0)),
                 ((2, 6, 7, 3), (0, 1, 0)), ((0, 4, 6, 2), (-1, 0, 0)), ((1, 3, 7, 5), (1, 0, 0))]
        for (a, b, c, d), nrm in quads:
            tris.append((nrm, v[a], v[b], v[c]))
            tris.append((nrm, v[a], v[c], v[d]))
    out = bytearray(b"urban canopy blocks".ljust(80, b" "))
    out += struct.pack("<I", len(tris))
    for nrm, a, b, c in tris:
        out += struct.pack("<12fH", *nrm, *a, *b, *c, 0)
    return bytes(out)


def urban_faces(nx: int = 320, ny: int = 160, nz: int = 96, dmin: float = 0.01):
    """Uniform in x and y; z clustered toward the ground by a partial cosine
    with first cell dmin (at the 96-cell grid; scaled by 96 / nz on others)."""
    Lx, Ly, Lz = URBAN_DOMAIN
    d0 = dmin * 96.0 / nz
    lo, hi = 1e-6, math.pi / 2 - 1e-6
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if _partial_cos(Lz, nz, mid)[1] < d0:
            lo = mid
        else:
            hi = mid
    return uniform_faces(0.0, Lx, nx), uniform_faces(0.0, Ly, ny), _partial_cos(Lz, nz, 0.5 * (lo + hi))


def log_law(z: float) -> float:
    """Inflow u(z) = (u* / kappa) ln((z + z0) / z0)."""
    return URBAN_USTAR / KARMAN * math.log((max(z, 0.0) + URBAN_Z0) / URBAN_Z0)


def setup_urban(s: I.Solver, nx: int = 320, ny: int = 160, nz: int = 96, Re: float = 2.0e4, dt: float = 0.001,
                tol: float = 1e-9, precision: int = I.PREC_FP64, turb: tuple | None = (I.TURBMODEL_VREMAN, 0.07),
                seed: int = URBAN_SEED):
    """The urban-canopy workload: West log-law inflow (VAL_FUNC), East
    pressure outlet, free-slip lateral and top boundaries, no-slip ground;
    blocks from urban_blocks(seed) through an in-memory STL; initial field the
    inflow profile everywhere."""
    xf, yf, zf = urban_faces(nx, ny, nz)
    s.set_grid(xf, yf, zf)
    s.set_params(Re, dt)
    zero = lambda t, x, y, z: 0.0
    s.set_bc(I.DIR_WEST, I.BC_VELOCITY_COMPONENT, I.VAL_FUNC, lambda t, x, y, z: log_law(z), zero, zero)
    s.set_bc(I.DIR_EAST, I.BC_PRESSURE, I.VAL_CONST, 0.0)
    s.set_bc(I.DIR_NORTH | I.DIR_SOUTH | I.DIR_UP, I.BC_FREE_SLIP_WALL)
    s.set_bc(I.DIR_DOWN, I.BC_STATIONARY_WALL)
    if turb is not None:
        s.set_turb_model(*turb)
    s.set_obstacle_stl_bytes(blocks_stl(urban_blocks(seed)))
    s.set_linear_solver(I.SOLVER_BiCGSTAB, I.PRECOND_AMG, tol)
    s.set_precision(precision)
    s.assemble()
    zc = cell_centres(zf)
    u = np.array([log_law(z) for z in zc])
    s.init_flow_const(0, 0, 0, 0)
    s.set("U1", np.broadcast_to(u[None, None, :], s.shape).copy())
    return s
