"""Convective outflow (BC_CONVECTIVE_OUTFLOW), an extension with no
counterpart in the reference.

The face velocity g obeys dg/dt + Uc dg/dn = 0, one explicit upwind step per
time step between the first interior cell centre and the face, before the
momentum solve; the pressure is held at 0 on the face. CPU tests pin the
oracle through what the scheme fixes, observed only through the solver's
outputs (the outer ghost layers, which carry the face value through the
Dirichlet interpolation rows):
  * a uniform stream is an exact steady solution;
  * the face values read back from the ghost layers obey the upwind relation
    with the time-n interior velocity, for a set speed and for the mean
    outward speed, on a high (x) and a low (y) face;
  * the projected face velocities balance inflow and outflow.
GPU tests compare the HIP path with the oracle on the same cases.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from test_oracle import bind_oracle_hooks, coords


def stretched(n, a, b, beta=0.3):
    s = np.linspace(0.0, 1.0, n + 1)
    return a + (b - a) * (s + beta * np.sin(math.pi * s) / math.pi * (1 - s) * 2) / 1.0


def channel(I, s, outlet, uc, n=(16, 10, 8), Re=50.0, dt=0.01, tol=1e-12, precision=None, perturb=True):
    """Box with inflow opposite the convective outlet and free-slip sides.
    outlet: "east" (flow +x) or "south" (flow -y)."""
    nx, ny, nz = n
    xf = stretched(nx, 0.0, 2.0)
    yf = stretched(ny, 0.0, 1.0, 0.2)
    zf = np.linspace(0.0, 0.8, nz + 1)
    s.set_grid(xf, yf, zf)
    s.set_params(Re, dt)
    if outlet == "east":
        s.set_bc(I.DIR_WEST, I.BC_VELOCITY_COMPONENT, I.VAL_CONST, 1.0, 0.0, 0.0)
        s.set_bc(I.DIR_EAST, I.BC_CONVECTIVE_OUTFLOW, uc)
        s.set_bc(I.DIR_NORTH | I.DIR_SOUTH | I.DIR_DOWN | I.DIR_UP, I.BC_FREE_SLIP_WALL)
    else:
        s.set_bc(I.DIR_NORTH, I.BC_VELOCITY_COMPONENT, I.VAL_CONST, 0.0, -1.0, 0.0)
        s.set_bc(I.DIR_SOUTH, I.BC_CONVECTIVE_OUTFLOW, uc)
        s.set_bc(I.DIR_EAST | I.DIR_WEST | I.DIR_DOWN | I.DIR_UP, I.BC_FREE_SLIP_WALL)
    s.set_obstacle(None)
    s.set_linear_solver(I.SOLVER_BiCGSTAB, I.PRECOND_AMG, tol)
    if precision is not None:
        s.set_precision(precision)
    s.assemble()
    if outlet == "east":
        base = (1.0, 0.0, 0.0)
    else:
        base = (0.0, -1.0, 0.0)
    if perturb:
        # a smooth swirl on top of the stream, so the outlet sees varying data
        s.init_flow_func(lambda x, y, z: base[0] + 0.2 * math.sin(2.1 * y + 0.5) * math.cos(1.3 * x),
                         lambda x, y, z: base[1] + 0.15 * math.cos(1.7 * x) * math.sin(2.0 * z + 0.3),
                         lambda x, y, z: 0.1 * math.sin(1.9 * x + 0.4) * math.cos(2.3 * y),
                         lambda x, y, z: 0.0)
    else:
        s.init_flow_const(*base, 0.0)
    return s


def face_values(s, outlet, widths):
    """Face values (3, n1, n2) over the interior tangential cells, read from
    the first outer ghost layer and the first interior cell through the
    Dirichlet interpolation row, plus the interior values themselves."""
    nx, ny, nz = s.Nx, s.Ny, s.Nz
    dx, dy = widths
    u = [s.get(n) for n in ("U1", "U2", "U3")]
    if outlet == "east":
        o, q = nx + 2, nx + 1                 # padded indices of i = Nx, Nx - 1
        ao, bi = dx[o] / 2, dx[q] / 2
        sl = lambda a, i: a[i, 2:ny + 2, 2:nz + 2]
    else:
        o, q = 1, 2                           # j = -1, 0
        ao, bi = dy[o] / 2, dy[q] / 2
        sl = lambda a, j: a[2:nx + 2, j, 2:nz + 2]
    g = np.stack([(bi * sl(a, o) + ao * sl(a, q)) / (ao + bi) for a in u])
    uq = np.stack([sl(a, q) for a in u])
    return g, uq, bi


@pytest.mark.parametrize("uc", [1.0, 0.0])
def test_convective_uniform_stream_is_steady(I, new_oracle, uc):
    s = channel(I, new_oracle(), "east", uc, perturb=False)
    s.iterate(4)
    u1, u2, u3, p = (s.get(n) for n in ("U1", "U2", "U3", "P"))
    inner = (slice(2, -2),) * 3
    assert np.max(np.abs(u1[inner] - 1.0)) < 1e-11
    assert np.max(np.abs(u2[inner])) < 1e-11 and np.max(np.abs(u3[inner])) < 1e-11
    assert np.max(np.abs(p[inner])) < 1e-11


@pytest.mark.parametrize("outlet", ["east", "south"])
@pytest.mark.parametrize("uc", [0.7, 0.0])
def test_convective_face_update_relation(I, new_oracle, outlet, uc):
    s = channel(I, new_oracle(), outlet, uc)
    bind_oracle_hooks(s.lib)
    _, _, _, dx, dy, dz = coords(s)
    s.iterate(1)
    g1, uq1, bi = face_values(s, outlet, (dx, dy))
    s.iterate(1)  # AB2 restarts, the face values carry over
    g2, _, _ = face_values(s, outlet, (dx, dy))
    nx, ny, nz = s.Nx, s.Ny, s.Nz
    if uc > 0:
        speed = uc
    else:
        # area-weighted mean outward normal velocity of g at time n
        if outlet == "east":
            A = np.outer(dy[2:ny + 2], dz[2:nz + 2])
            speed = max(float(np.sum(g1[0] * A) / np.sum(A)), 0.0)
        else:
            A = np.outer(dx[2:nx + 2], dz[2:nz + 2])
            speed = max(float(np.sum(-g1[1] * A) / np.sum(A)), 0.0)
        assert 0.5 < speed < 1.5
    want = g1 - s.dt * speed * (g1 - uq1) / bi
    # the face values must actually move, or the relation pins nothing
    assert np.max(np.abs(g2 - g1)) > 1e-4
    np.testing.assert_allclose(g2, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("outlet", ["east", "south"])
def test_convective_mass_balance(I, new_oracle, outlet):
    s = channel(I, new_oracle(), outlet, 0.0)
    bind_oracle_hooks(s.lib)
    _, _, _, dx, dy, dz = coords(s)
    s.iterate(3)
    nx, ny, nz = s.Nx, s.Ny, s.Nz
    if outlet == "east":
        U = s.get("FACE_U1")
        A = np.outer(dy[2:ny + 2], dz[2:nz + 2])
        qin = np.sum(U[0, 2:ny + 2, 2:nz + 2] * A)
        qout = np.sum(U[nx, 2:ny + 2, 2:nz + 2] * A)
    else:
        U = s.get("FACE_U2")
        A = np.outer(dx[2:nx + 2], dz[2:nz + 2])
        qin = -np.sum(U[2:nx + 2, ny, 2:nz + 2] * A)
        qout = -np.sum(U[2:nx + 2, 0, 2:nz + 2] * A)
    assert abs(qin - float(np.sum(A))) < 1e-12 * np.sum(A)   # inflow = 1 x area
    assert abs(qout - qin) < 1e-9 * abs(qin)
    assert s.max_div() < 1e-9


# ---------------------------------------------------------------------------
# HIP path against the oracle
# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("outlet", ["east", "south"])
@pytest.mark.parametrize("uc", [0.7, 0.0])
def test_gpu_convective_matches_oracle(I, new_oracle, new_gpu, outlet, uc):
    o = channel(I, new_oracle(), outlet, uc, n=(24, 16, 12))
    g = channel(I, new_gpu(), outlet, uc, n=(24, 16, 12))
    o.iterate(5)
    g.iterate(5)
    for name in ("U1", "U2", "U3", "P"):
        a, b = o.get(name), g.get(name)
        scale = max(np.max(np.abs(a)), 1e-30)
        assert np.max(np.abs(a - b)) / scale < 1e-8, name
    assert g.max_div() < 1e-9


@pytest.mark.gpu
def test_gpu_convective_sphere_mixed(I, new_oracle, new_gpu, cases):
    """Sphere in a stream, convective outlet, FP64 and mixed precision."""
    def setup(s, prec):
        xf, yf, zf = cases.sphere_faces(48, 32, 32, cases.SPHERE_DMIN * 128.0 / 32)
        s.set_grid(xf, yf, zf)
        s.set_params(300.0, 0.002)
        s.set_bc(I.DIR_WEST, I.BC_VELOCITY_COMPONENT, I.VAL_CONST, 1.0, 0.0, 0.0)
        s.set_bc(I.DIR_EAST, I.BC_CONVECTIVE_OUTFLOW, 0.0)
        s.set_bc(I.DIR_NORTH | I.DIR_SOUTH | I.DIR_DOWN | I.DIR_UP, I.BC_FREE_SLIP_WALL)
        s.set_obstacle_sphere(0.0, 0.0, 0.0, 0.5)
        s.set_linear_solver(I.SOLVER_BiCGSTAB, I.PRECOND_AMG, 1e-11)
        if prec is not None:
            s.set_precision(prec)
        s.assemble()
        s.init_flow_const(1.0, 0.0, 0.0, 0.0)
        return s
    o = setup(new_oracle(), None)
    o.iterate(4)
    ref = [o.get(n) for n in ("U1", "U2", "U3")]
    fl = o.flags() == I.FLAG_FLUID
    for prec in (I.PREC_FP64, I.PREC_MIXED_KRYLOV):
        g = setup(new_gpu(), prec)
        g.iterate(4)
        for a, name in zip(ref, ("U1", "U2", "U3")):
            b = g.get(name)
            rel = np.linalg.norm((a - b)[fl]) / np.linalg.norm(a[fl])
            assert rel < 1e-7, (prec, name, rel)
        assert g.max_div() < 1e-8
