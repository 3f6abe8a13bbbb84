"""ctypes binding of the IBMSolver C API (include/ibm3d.h, include/ibm3d_ext.h).

The same binding drives any shared library that exports the reference's
``IBMSolver_*`` API: the MI355X library ``libibm3d_hip.so`` built from
``csrc/`` (the product) and, in the tests, the CPU oracle. Function names,
argument order and error behaviour are the reference's (src/ibm3d_setup.h,
src/ibm3d_init.h, src/ibm3d_fracstep.h, src/ibm3d_export.h); arrays come back
as numpy views in the reference's padded host layout (src/utils.h:8-16).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Callable, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libibm3d_hip.so")

# enums: src/ibm3d.h:16-98
SOLVER_AMG, SOLVER_PCG, SOLVER_BiCGSTAB, SOLVER_GMRES = range(4)
PRECOND_NONE, PRECOND_AMG = range(2)
DIR_NORTH, DIR_EAST, DIR_SOUTH, DIR_WEST, DIR_DOWN, DIR_UP = (1 << i for i in range(6))
(BC_VELOCITY_COMPONENT, BC_PRESSURE, BC_STATIONARY_WALL, BC_FREE_SLIP_WALL,
 BC_SYMMETRIC, BC_ALL_PERIODIC, BC_VELOCITY_PERIODIC) = range(7)
BC_CONVECTIVE_OUTFLOW = 7  # extension: convective outflow, varargs (double Uc)
VAL_CONST, VAL_FUNC = range(2)
TURBMODEL_NONE, TURBMODEL_SMAGORINSKY, TURBMODEL_WALE, TURBMODEL_VREMAN, TURBMODEL_SMAGORINSKY_VD = range(5)
FLAG_FLUID, FLAG_GHOST, FLAG_SOLID = 1, 2, 3

# include/ibm3d_ext.h
FIELDS = ["U1", "U2", "U3", "P", "U1_STAR", "U2_STAR", "U3_STAR", "P_PRIME",
          "N1", "N2", "N3", "LVSET", "FACE_U1", "FACE_U2", "FACE_U3",
          "TAU11", "TAU12", "TAU13", "TAU22", "TAU23", "TAU33"]
FIELD = {name: i for i, name in enumerate(FIELDS)}
PREC_FP64, PREC_MIXED, PREC_MIXED_KRYLOV = 0, 1, 2

INIT_FUNC = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, C.c_double)


# include/ibm3d_serve.h
class ServeInit(C.Structure):
    _fields_ = [("u1", C.c_double), ("u2", C.c_double), ("u3", C.c_double), ("p", C.c_double),
                ("amp", C.c_double), ("seed", C.c_ulonglong)]


class ServeResult(C.Structure):
    _fields_ = [("state", C.c_int), ("t", C.c_double * 3), ("iters", C.c_int * 4), ("max_div", C.c_double),
                ("device_bytes", C.c_double), ("error", C.c_char * 256)]
BC_FUNC = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, C.c_double, C.c_double)

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

_SIGS = {
    "IBMSolver_new": (_vp, [C.c_int] * 5),
    "IBMSolver_destroy": (None, [_vp]),
    "IBMSolver_set_grid": (None, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]),
    "IBMSolver_set_params": (None, [_vp, C.c_double, C.c_double]),
    "IBMSolver_set_obstacle": (None, [_vp, _vp]),
    "IBMSolver_set_linear_solver": (None, [_vp, C.c_int, C.c_int, C.c_double]),
    "IBMSolver_set_autosave": (None, [_vp, C.c_char_p, C.c_int]),
    "IBMSolver_assemble": (None, [_vp]),
    "IBMSolver_init_flow_const": (None, [_vp, C.c_double, C.c_double, C.c_double, C.c_double]),
    "IBMSolver_init_flow_func": (None, [_vp, _vp, _vp, _vp, _vp]),
    "IBMSolver_init_flow_file": (None, [_vp, C.c_char_p]),
    "IBMSolver_iterate": (None, [_vp, C.c_int, C.c_bool]),
    "IBMSolver_calc_tau_r": (None, [_vp]),
    "IBMSolver_export_result": (None, [_vp, C.c_char_p]),
    "IBMSolver_export_lvset_flag": (None, [_vp, C.c_char_p]),
    "IBMSolver_export_intermediate": (None, [_vp, C.c_char_p]),
    "IBMSolver_ghost_interp": (None, [_vp, C.c_int, C.c_int, C.c_int, _ip, _dp]),
    "IBMSolver_set_obstacle_sphere": (None, [_vp] + [C.c_double] * 4),
    "IBMSolver_set_precision": (None, [_vp, C.c_int]),
    "IBMSolver_set_mg_params": (None, [_vp, C.c_double, C.c_int, C.c_int, C.c_int]),
    "IBMSolver_get_field": (None, [_vp, C.c_int, _dp]),
    "IBMSolver_set_field": (None, [_vp, C.c_int, _dp]),
    "IBMSolver_get_flag": (None, [_vp, _ip]),
    "IBMSolver_get_iter": (C.c_int, [_vp]),
    "IBMSolver_get_time": (C.c_double, [_vp]),
    "IBMSolver_get_last_norms": (None, [_vp, _dp]),
    "IBMSolver_get_last_iters": (None, [_vp, _ip]),
    "IBMSolver_get_ghost_cells": (C.c_int, [_vp, _ip]),
    "IBMSolver_max_div": (C.c_double, [_vp]),
    "Polyhedron_new": (_vp, []),
    "Polyhedron_destroy": (None, [_vp]),
    "Polyhedron_read_stl": (None, [_vp, _vp]),
}
# variadic: IBMSolver_set_bc, IBMSolver_set_ext_force, IBMSolver_set_turb_model

# optional symbols (present in the product, absent in the oracle or vice versa)
_OPTIONAL = {
    "IBMSolver_calc_body_force": (None, [_vp, _dp]),
    "IBMSolver_set_force_history": (None, [_vp, C.c_int]),
    "IBMSolver_set_momentum_solver": (None, [_vp, C.c_int]),
    "IBMSolver_get_force_history": (C.c_int, [_vp, _dp, C.c_int]),
    "IBMSolver_enable_timing": (None, [_vp, C.c_int]),
    "IBMSolver_set_graph_replay": (None, [_vp, C.c_int]),
    "IBMSolver_get_step_times": (C.c_int, [_vp, _dp, C.c_int]),
    "IBMSolver_reset_timing": (None, [_vp]),
    "IBMSolver_get_kernel_stat": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int, _dp,
                                            C.POINTER(C.c_long), _dp]),
    "IBMSolver_get_device_bytes": (C.c_double, [_vp]),
    "IBMSolver_get_mg_levels": (C.c_int, [_vp, _ip, C.c_int]),
    "IBMSolver_set_mg_light_levels": (None, [_vp, C.c_int]),
    "IBMSolver_set_mg_relaxation": (None, [_vp, C.c_double]),
    "IBMSolver_get_pressure_mg_bits": (C.c_int, [_vp]),
    "IBMSolver_estimate_device_bytes": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "IBMServer_new": (_vp, [C.c_double, C.c_int]),
    "IBMServer_submit": (C.c_int, [_vp, C.c_double, C.c_int, C.c_void_p, C.c_void_p, _vp]),
    "IBMServer_submit_solver": (C.c_int, [_vp, _vp, C.c_double, C.c_int, C.c_void_p, C.c_void_p]),
    "IBMServer_wait": (None, [_vp]),
    "IBMServer_job_times": (C.c_int, [_vp, C.c_int, _dp]),
    "IBMServer_job_result": (C.c_int, [_vp, C.c_int, C.c_void_p]),
    "IBMServer_failed_jobs": (C.c_int, [_vp]),
    "IBMServer_peak_bytes": (C.c_double, [_vp]),
    "IBMServer_destroy": (None, [_vp]),
}

_LIBS: dict[str, C.CDLL] = {}


def load(path: str | None = None) -> C.CDLL:
    """Loads and binds an IBMSolver library; raises if it is missing."""
    path = os.path.abspath(path or LIB_PATH)
    if path in _LIBS:
        return _LIBS[path]
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is not built: run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = C.CDLL(path, mode=C.RTLD_LOCAL)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    for name, (res, args) in _OPTIONAL.items():
        if hasattr(lib, name):
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
    for name in ("IBMSolver_set_bc", "IBMSolver_set_ext_force", "IBMSolver_set_turb_model"):
        getattr(lib, name).restype = None
    _LIBS[path] = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class Solver:
    """Thin object wrapper over one IBMSolver handle of a given library."""

    def __init__(self, lib: C.CDLL | str | None = None):
        self.lib = lib if isinstance(lib, C.CDLL) else load(lib)
        self.h = self.lib.IBMSolver_new(1, 0, 1, 1, 1)
        if not self.h:
            raise RuntimeError("IBMSolver_new failed")
        self._keep: list = []  # callbacks must outlive the solver
        self.shape = None

    # --- setup -----------------------------------------------------------
    def set_grid(self, xf: Sequence[float], yf: Sequence[float], zf: Sequence[float]):
        xf, yf, zf = (np.ascontiguousarray(a, dtype=np.float64) for a in (xf, yf, zf))
        self.Nx, self.Ny, self.Nz = len(xf) - 1, len(yf) - 1, len(zf) - 1
        self.xf, self.yf, self.zf = xf, yf, zf
        self.shape = (self.Nx + 4, self.Ny + 4, self.Nz + 4)
        self.lib.IBMSolver_set_grid(self.h, self.Nx, self.Ny, self.Nz, _dptr(xf), _dptr(yf), _dptr(zf))

    def set_params(self, Re: float, dt: float):
        self.Re, self.dt = float(Re), float(dt)
        self.lib.IBMSolver_set_params(self.h, float(Re), float(dt))

    def set_bc(self, direction: int, bctype: int, *vals):
        """vals as in src/ibm3d_setup.c:378-415, e.g. (VAL_CONST, 1., 0., 0.);
        BC_CONVECTIVE_OUTFLOW takes (Uc,), Uc <= 0 for the mean outward
        velocity."""
        if bctype == BC_CONVECTIVE_OUTFLOW:
            uc = float(vals[0]) if vals else 0.0
            self.lib.IBMSolver_set_bc(C.c_void_p(self.h), C.c_int(direction), C.c_int(bctype), C.c_double(uc))
            return
        args = []
        for v in vals:
            if isinstance(v, int) and not isinstance(v, bool) and not args:
                args.append(C.c_int(v))      # IBMSolverValType
            elif callable(v):
                cb = BC_FUNC(v)
                self._keep.append(cb)
                args.append(C.cast(cb, _vp))
            elif isinstance(v, C._CFuncPtr):
                args.append(C.cast(v, _vp))
            else:
                args.append(C.c_double(float(v)))
        self.lib.IBMSolver_set_bc(C.c_void_p(self.h), C.c_int(direction), C.c_int(bctype), *args)

    def set_ext_force(self, f1, f2, f3):
        if all(isinstance(v, (int, float)) for v in (f1, f2, f3)):
            self.lib.IBMSolver_set_ext_force(C.c_void_p(self.h), C.c_int(VAL_CONST), C.c_double(f1),
                                             C.c_double(f2), C.c_double(f3))
        else:
            cbs = [BC_FUNC(f) for f in (f1, f2, f3)]
            self._keep += cbs
            self.lib.IBMSolver_set_ext_force(C.c_void_p(self.h), C.c_int(VAL_FUNC), *[C.cast(c, _vp) for c in cbs])

    def set_turb_model(self, model: int, coef: float = 0.0, aplus: float = 26.0, u_tau: float = 1.0,
                       walls: int = 0):
        """coef: Cs (Smagorinsky) or c (Vreman); aplus, u_tau, walls: van
        Driest damping of TURBMODEL_SMAGORINSKY_VD."""
        if model == TURBMODEL_NONE:
            self.lib.IBMSolver_set_turb_model(C.c_void_p(self.h), C.c_int(model))
        elif model == TURBMODEL_SMAGORINSKY_VD:
            self.lib.IBMSolver_set_turb_model(C.c_void_p(self.h), C.c_int(model), C.c_double(coef),
                                              C.c_double(aplus), C.c_double(u_tau), C.c_int(walls))
        else:
            self.lib.IBMSolver_set_turb_model(C.c_void_p(self.h), C.c_int(model), C.c_double(coef))

    def set_obstacle_sphere(self, cx: float, cy: float, cz: float, r: float):
        self.lib.IBMSolver_set_obstacle_sphere(self.h, cx, cy, cz, r)

    def set_obstacle(self, poly):
        self.lib.IBMSolver_set_obstacle(self.h, poly)

    def set_obstacle_stl(self, path: str):
        """Reads an STL file with Polyhedron_read_stl and sets it as the
        obstacle; the polyhedron is destroyed after assemble (it is borrowed
        until then, as in the reference)."""
        import ctypes.util
        libc = C.CDLL(ctypes.util.find_library("c"))
        libc.fopen.restype = C.c_void_p
        libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
        libc.fclose.argtypes = [C.c_void_p]
        fp = libc.fopen(os.fsencode(path), b"rb")
        if not fp:
            raise FileNotFoundError(path)
        poly = self.lib.Polyhedron_new()
        self.lib.Polyhedron_read_stl(poly, fp)
        libc.fclose(fp)
        self._poly = poly
        self.lib.IBMSolver_set_obstacle(self.h, poly)

    def set_obstacle_stl_bytes(self, data: bytes):
        """Reads a binary STL held in memory through Polyhedron_read_stl (an
        in-memory FILE* from fmemopen, so the reference's reader API is used
        unchanged) and sets it as the obstacle, borrowed until assemble."""
        import ctypes.util
        libc = C.CDLL(ctypes.util.find_library("c"))
        libc.fmemopen.restype = C.c_void_p
        libc.fmemopen.argtypes = [C.c_void_p, C.c_size_t, C.c_char_p]
        libc.fclose.argtypes = [C.c_void_p]
        buf = C.create_string_buffer(bytes(data), len(data))
        fp = libc.fmemopen(buf, len(data), b"rb")
        if not fp:
            raise OSError("fmemopen failed")
        poly = self.lib.Polyhedron_new()
        self.lib.Polyhedron_read_stl(poly, fp)
        libc.fclose(fp)
        self._poly = poly
        self.lib.IBMSolver_set_obstacle(self.h, poly)

    def set_linear_solver(self, solver=SOLVER_BiCGSTAB, precond=PRECOND_AMG, tol=1e-10):
        self.lib.IBMSolver_set_linear_solver(self.h, solver, precond, tol)

    def set_precision(self, mode: int):
        self.lib.IBMSolver_set_precision(self.h, mode)

    def set_mg_params(self, gamma: float = 3.5, pre: int = 2, post: int = 2, smoother: int = 1):
        self.lib.IBMSolver_set_mg_params(self.h, gamma, pre, post, smoother)

    def set_mg_relaxation(self, omega: float = 1.3):
        """Over-relaxation factor of the pressure multigrid smoother."""
        self.lib.IBMSolver_set_mg_relaxation(self.h, float(omega))

    def pressure_mg_bits(self) -> int:
        """32 or 64: the precision of the pressure multigrid cycle in use."""
        return int(self.lib.IBMSolver_get_pressure_mg_bits(self.h))

    def set_mg_light_levels(self, from_level: int = 1):
        if hasattr(self.lib, "IBMSolver_set_mg_light_levels"):
            self.lib.IBMSolver_set_mg_light_levels(self.h, int(from_level))

    def assemble(self):
        self.lib.IBMSolver_assemble(self.h)
        poly = getattr(self, "_poly", None)
        if poly:
            self.lib.Polyhedron_destroy(poly)
            self._poly = None

    # --- init / run --------------------------------------------------------
    def init_flow_const(self, u1=0.0, u2=0.0, u3=0.0, p=0.0):
        self.lib.IBMSolver_init_flow_const(self.h, u1, u2, u3, p)

    def init_flow_func(self, f1: Callable, f2: Callable, f3: Callable, fp: Callable):
        cbs = [INIT_FUNC(f) for f in (f1, f2, f3, fp)]
        self._keep += cbs
        self.lib.IBMSolver_init_flow_func(self.h, *[C.cast(c, _vp) for c in cbs])

    def iterate(self, steps: int, verbose: bool = False):
        self.lib.IBMSolver_iterate(self.h, int(steps), bool(verbose))

    # --- readback ----------------------------------------------------------
    def field_shape(self, name: str):
        nx, ny, nz = self.Nx, self.Ny, self.Nz
        if name == "FACE_U1":
            return (nx + 1, ny + 4, nz + 4)
        if name == "FACE_U2":
            return (nx + 4, ny + 1, nz + 4)
        if name == "FACE_U3":
            return (nx + 4, ny + 4, nz + 1)
        return self.shape

    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.field_shape(name), dtype=np.float64)
        self.lib.IBMSolver_get_field(self.h, FIELD[name], _dptr(out))
        return out

    def set(self, name: str, arr: np.ndarray):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        assert arr.shape == self.shape
        self.lib.IBMSolver_set_field(self.h, FIELD[name], _dptr(arr))

    def flags(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.int32)
        self.lib.IBMSolver_get_flag(self.h, out.ctypes.data_as(_ip))
        return out

    def ghost_cells(self) -> np.ndarray:
        n = self.lib.IBMSolver_get_ghost_cells(self.h, None)
        out = np.empty((n, 3), dtype=np.int32)
        if n:
            self.lib.IBMSolver_get_ghost_cells(self.h, out.ctypes.data_as(_ip))
        return out

    def last_norms(self) -> np.ndarray:
        out = np.zeros(4)
        self.lib.IBMSolver_get_last_norms(self.h, _dptr(out))
        return out

    def last_iters(self) -> np.ndarray:
        out = np.zeros(4, dtype=np.int32)
        self.lib.IBMSolver_get_last_iters(self.h, out.ctypes.data_as(_ip))
        return out

    @property
    def iter(self) -> int:
        return self.lib.IBMSolver_get_iter(self.h)

    @property
    def time(self) -> float:
        return self.lib.IBMSolver_get_time(self.h)

    def max_div(self) -> float:
        return self.lib.IBMSolver_max_div(self.h)

    def body_force(self) -> np.ndarray:
        out = np.zeros(3)
        self.lib.IBMSolver_calc_body_force(self.h, _dptr(out))
        return out

    # --- timing (product only) ----------------------------------------------
    def set_momentum_solver(self, mode: int):
        self.lib.IBMSolver_set_momentum_solver(self.h, int(mode))

    def set_force_history(self, max_steps: int):
        self.lib.IBMSolver_set_force_history(self.h, int(max_steps))

    def force_history(self) -> np.ndarray:
        n = self.lib.IBMSolver_get_force_history(self.h, None, 0)
        out = np.zeros((n, 3))
        if n:
            self.lib.IBMSolver_get_force_history(self.h, _dptr(out), n)
        return out

    def enable_timing(self, mode: int = 1):
        """0 off, 1 every kernel (multigrid cycles eager), 2 kernels outside
        the cycles plus whole graph replays of the cycles."""
        self.lib.IBMSolver_enable_timing(self.h, int(mode))

    def set_graph_replay(self, on: bool):
        self.lib.IBMSolver_set_graph_replay(self.h, int(on))

    def reset_timing(self):
        self.lib.IBMSolver_reset_timing(self.h)

    def step_times(self) -> np.ndarray:
        n = self.lib.IBMSolver_get_step_times(self.h, None, 0)
        out = np.zeros(n)
        if n:
            self.lib.IBMSolver_get_step_times(self.h, _dptr(out), n)
        return out

    def kernel_stats(self) -> list[dict]:
        n = self.lib.IBMSolver_get_kernel_stat(self.h, -1, None, 0, None, None, None)
        out = []
        for i in range(n):
            name = C.create_string_buffer(64)
            ms, cnt, nbytes = C.c_double(), C.c_long(), C.c_double()
            self.lib.IBMSolver_get_kernel_stat(self.h, i, name, 64, C.byref(ms), C.byref(cnt),
                                               C.byref(nbytes))
            out.append({"name": name.value.decode(), "ms": ms.value, "count": cnt.value,
                        "bytes": nbytes.value})
        return out

    def mg_levels(self) -> list[tuple[int, int, int]]:
        n = self.lib.IBMSolver_get_mg_levels(self.h, None, 0)
        buf = (C.c_int * (3 * max(n, 1)))()
        self.lib.IBMSolver_get_mg_levels(self.h, buf, n)
        return [tuple(buf[3 * l:3 * l + 3]) for l in range(n)]

    def device_bytes(self) -> float:
        return self.lib.IBMSolver_get_device_bytes(self.h)

    def detach(self):
        """Hands the handle over to a caller that will destroy it (e.g. the
        server); this object no longer owns it."""
        h, self.h = self.h, None
        return h

    def close(self):
        if self.h:
            self.lib.IBMSolver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
