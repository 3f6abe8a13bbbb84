"""The ibm3d.in driver (drivers/ibm3d_main.c, after src/main.c:61-173).

CPU: it builds against the C-ABI headers and library with -Wall -Werror, and
refuses a missing input file. GPU: it runs the reference's own input sample
(tests/golden/ibm3d.in: STL sphere, 192x48x48, Re 200, dt 0.01) for a few
steps; the netCDF result and level-set/flag files are read back and compared
with the same case built through the Python binding (flags bit-exact, fields
equal to float32 rounding)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVERS = os.path.join(ROOT, "2020researchproject_amd", "drivers")
EXE = os.path.join(DRIVERS, "ibm3d_main")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _read_sample():
    toks = open(os.path.join(GOLDEN, "ibm3d.in")).read().split()
    it = iter(toks)
    out = {}
    assert next(it) == "STLFile"
    out["stl"] = next(it)
    for ax in ("Nx", "Ny", "Nz"):
        assert next(it) == ax
        n = int(next(it))
        out[ax] = np.array([float(next(it)) for _ in range(n + 1)])
    assert next(it) == "Re"
    out["Re"] = float(next(it))
    assert next(it) == "dt"
    out["dt"] = float(next(it))
    assert next(it) == "numtstep"
    out["steps"] = int(next(it))
    return out


def _write_input(path, stl, out_base, init=0, init_file="none"):
    s = _read_sample()
    with open(path, "w") as f:
        f.write(f"STLFile     {stl}\n\n")
        for ax in ("Nx", "Ny", "Nz"):
            f.write(f"{ax}          {len(s[ax]) - 1}\n" + " ".join(f"{v:.8f}" for v in s[ax]) + "\n")
        f.write(f"\nRe          {s['Re']:g}\ndt          {s['dt']:g}\nnumtstep    {s['steps']}\n\n")
        f.write(f"InitUsingFile       {init}\nInitFile            {init_file}\nOutputFile          {out_base}\n")
    return s


def test_driver_builds_and_rejects_missing_input(tmp_path):
    subprocess.run(["make", "-C", DRIVERS, "-B", "CC=gcc -Werror"], check=True, capture_output=True)
    assert os.access(EXE, os.X_OK)
    r = subprocess.run([EXE, str(tmp_path / "nope.in")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr
    # an unknown precision is refused before anything is read
    r = subprocess.run([EXE, str(tmp_path / "nope.in"), "-precision", "fp16"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "unknown precision 'fp16'" in r.stderr


def test_sample_input_parses():
    s = _read_sample()
    assert (len(s["Nx"]) - 1, len(s["Ny"]) - 1, len(s["Nz"]) - 1) == (192, 48, 48)
    assert s["Re"] == 200 and s["dt"] == 0.01 and s["steps"] == 100


@pytest.mark.gpu
def test_driver_runs_reference_sample(tmp_path, I):
    from scipy.io import netcdf_file
    base = str(tmp_path / "sphere")
    inp = str(tmp_path / "ibm3d.in")
    s = _write_input(inp, os.path.join(GOLDEN, "sphere.stl"), base)
    r = subprocess.run([EXE, inp, "-steps", "3"], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    # the same case through the Python binding
    g = I.Solver()
    g.set_grid(s["Nx"], s["Ny"], s["Nz"])
    g.set_params(s["Re"], s["dt"])
    g.set_bc(I.DIR_WEST, I.BC_VELOCITY_COMPONENT, I.VAL_CONST, 1.0, 0.0, 0.0)
    g.set_bc(I.DIR_EAST, I.BC_PRESSURE, I.VAL_CONST, 0.0)
    g.set_bc(I.DIR_NORTH | I.DIR_SOUTH | I.DIR_DOWN | I.DIR_UP, I.BC_FREE_SLIP_WALL)
    g.set_obstacle_stl(os.path.join(GOLDEN, "sphere.stl"))
    g.set_linear_solver(I.SOLVER_BiCGSTAB, I.PRECOND_AMG, 1e-6)
    g.assemble()
    g.init_flow_const(1.0, 0.0, 0.0, 0.0)
    g.iterate(3)
    with netcdf_file(base + "_lvset_flag.nc", "r", mmap=False) as nc:
        assert np.array_equal(nc.variables["flag"][:].transpose(2, 1, 0), g.flags())
    with netcdf_file(base + ".nc", "r", mmap=False) as nc:
        assert int(nc.variables["iter"].getValue()) == 3
        u = nc.variables["u"][0].transpose(2, 1, 0)
        np.testing.assert_array_equal(u, g.get("U1").astype(np.float32))
    g.close()


SERVE = os.path.join(DRIVERS, "ibm3d_serve")


def _serve(tmp_path, lines, budget_gib=4, slots=2):
    cf = tmp_path / "cases.txt"
    cf.write_text("\n".join(lines) + "\n")
    r = subprocess.run([SERVE, str(budget_gib), str(slots), str(cf)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-1000:], flush=True)
    jobs = {}
    for line in r.stdout.splitlines():
        if line.startswith("job "):
            t = line.split()
            jobs[int(t[1])] = {"kind": t[2], "state": t[4], "iters": [int(v) for v in t[12:16]],
                               "max_div": float(t[17]), "error": line.split('error "', 1)[1].rstrip('"')}
    return r, jobs


def test_serve_driver_failed_jobs_fail_alone(tmp_path):
    """The C serving driver (drivers/ibm3d_serve.c: cases configured in C and
    submitted with IBMServer_submit_solver, no Python): set-ups that
    IBMSolver_assemble rejects fail as jobs with the reference's message, the
    process goes on and exits 0. No GPU needed: they fail before any device
    work."""
    subprocess.run(["make", "-C", DRIVERS, "-B", "CC=gcc -Werror"], check=True, capture_output=True)
    r, jobs = _serve(tmp_path, ["badperiodic 8 3", "badperiodic 10 2"])
    assert r.returncode == 0
    assert len(jobs) == 2
    for j in jobs.values():
        assert j["state"] == "failed" and j["error"] == "Inconsistent periodic boundary condition"
    assert "jobs 2 failed 2" in r.stdout


@pytest.mark.parametrize("line, message", [("cavity 16 5", "malformed cavity case"),
                                           ("sphere 48 24 24 6", "malformed sphere case"),
                                           ("tunnel 16 5", "unknown case kind 'tunnel'")])
def test_serve_driver_malformed_input_is_an_error(tmp_path, line, message):
    """A case line the C serving driver cannot read ends the reading with an
    error on stderr and exit status 2; the job read before it still runs (here
    a set-up that fails before any device work) and is reported."""
    subprocess.run(["make", "-C", DRIVERS, "CC=gcc -Werror"], check=True, capture_output=True)
    r, jobs = _serve(tmp_path, ["badperiodic 8 3", line, "badperiodic 10 2"])
    assert r.returncode == 2
    assert message in r.stderr
    assert len(jobs) == 1 and next(iter(jobs.values()))["state"] == "failed"
    assert "jobs 1 failed 1" in r.stdout


@pytest.mark.gpu
def test_serve_driver_runs_a_mix(tmp_path, I, cases):
    """The C serving driver on the GPU: two cavities and a sphere run beside a
    rejected set-up, which fails alone. Each cavity's last-step iteration
    counts equal those of the same case built through the Python binding and
    run in this process (the native initial perturbation is the Python one)."""
    r, jobs = _serve(tmp_path, ["cavity 24 10 400", "badperiodic 16 5", "sphere 48 24 24 6 300",
                                "cavity 32 8 900"], budget_gib=8, slots=3)
    assert r.returncode == 0 and len(jobs) == 4
    assert jobs[1]["state"] == "failed" and "Inconsistent periodic" in jobs[1]["error"]
    for j in (0, 2, 3):
        assert jobs[j]["state"] == "done" and jobs[j]["max_div"] < 1e-8, jobs[j]
    for j, (n, steps, Re) in ((0, (24, 10, 400.0)), (3, (32, 8, 900.0))):
        s = I.Solver()
        cases.setup_cavity(s, n=n, Re=Re, tol=1e-10, precision=I.PREC_MIXED_KRYLOV)
        s.iterate(steps)
        assert s.last_iters().tolist() == jobs[j]["iters"], (j, jobs[j])
        assert s.max_div() == pytest.approx(jobs[j]["max_div"], rel=1e-3)
        s.close()
