"""The C-ABI library loads on a CPU-only host and exports every function the
public headers declare (no GPU calls)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import sys

from conftest import ROOT

HEADERS = ["ibm3d.h", "ibm3d_ext.h", "geo3d.h", "ibm3d_serve.h"]


def declared_functions():
    names = set()
    for h in HEADERS:
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b((?:IBMSolver|IBMServer|Polyhedron|Vector)_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_headers_declare_reference_api():
    names = declared_functions()
    # the reference's public entry points (src/ibm3d_setup.h:6-31, ibm3d_init.h:6-8,
    # ibm3d_fracstep.h:6, ibm3d_turb.h:6, ibm3d_export.h:6-9, geo3d.h:17-45)
    ref = {"IBMSolver_new", "IBMSolver_destroy", "IBMSolver_set_grid", "IBMSolver_set_params",
           "IBMSolver_set_ext_force", "IBMSolver_set_turb_model", "IBMSolver_set_bc", "IBMSolver_set_obstacle",
           "IBMSolver_set_linear_solver", "IBMSolver_set_autosave", "IBMSolver_assemble",
           "IBMSolver_ghost_interp", "IBMSolver_init_flow_const", "IBMSolver_init_flow_file",
           "IBMSolver_init_flow_func", "IBMSolver_iterate", "IBMSolver_calc_tau_r", "IBMSolver_export_result",
           "IBMSolver_export_lvset_flag", "IBMSolver_export_intermediate", "Polyhedron_new",
           "Polyhedron_destroy", "Polyhedron_read_stl", "Polyhedron_print_stats", "Polyhedron_scale",
           "Polyhedron_cpt", "Vector_add", "Vector_sub", "Vector_dot", "Vector_crs", "Vector_norm",
           "Vector_dist", "Vector_angle", "Vector_normalize", "Vector_lincom", "Vector_rot"}
    assert ref <= names


def test_library_exports_every_declared_symbol(gpu_lib):
    missing = [n for n in sorted(declared_functions()) if not hasattr(gpu_lib, n)]
    assert not missing, missing


def test_enum_values_match_reference(I):
    # src/ibm3d.h:16-98
    assert (I.DIR_NORTH, I.DIR_EAST, I.DIR_SOUTH, I.DIR_WEST, I.DIR_DOWN, I.DIR_UP) == (1, 2, 4, 8, 16, 32)
    assert (I.BC_VELOCITY_COMPONENT, I.BC_PRESSURE, I.BC_STATIONARY_WALL, I.BC_FREE_SLIP_WALL,
            I.BC_SYMMETRIC, I.BC_ALL_PERIODIC, I.BC_VELOCITY_PERIODIC) == tuple(range(7))
    assert (I.FLAG_FLUID, I.FLAG_GHOST, I.FLAG_SOLID) == (1, 2, 3)
    assert (I.SOLVER_AMG, I.SOLVER_PCG, I.SOLVER_BiCGSTAB, I.SOLVER_GMRES) == (0, 1, 2, 3)


def test_headers_compile_as_c_and_cxx(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "ibm3d.h"\nint main(void){IBMSolver *s = 0; (void)s; return DIR_UP;}\n')
    inc = os.path.join(ROOT, "include")
    for cc, ext in (("gcc", "c"), ("g++", "cpp")):
        f = tmp_path / f"t.{ext}"
        f.write_text(src.read_text())
        subprocess.run([cc, "-fsyntax-only", "-Wall", "-Werror", "-I", inc, str(f)], check=True)


def test_multi_gpu_request_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
