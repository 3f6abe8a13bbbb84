"""Host emulation of the fused red-black Gauss-Seidel kernel.

Compiles csrc/rbgs_fused.h for the CPU (tests/emul/rbgs_fused_emul.cpp: one
thread per lane, LDS as shared statics, barriers for __syncthreads) with
AddressSanitizer and UBSan, and checks one fused out-of-place sweep bit for
bit against a direct colour-0-then-colour-1 sweep over random operators,
flags (fluid / solid / ghost), pinned cells, periodic axes, zero input and
tile / plane-chunk seams, in FP64 and FP32."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "2020researchproject_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the ROCm headers")
def test_fused_rbgs_matches_two_colour_sweep(tmp_path):
    exe = str(tmp_path / "rbgs_emul")
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + CSRC,
           os.path.join(HERE, "emul", "rbgs_fused_emul.cpp"), "-o", exe, "-pthread"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    res = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert res.returncode == 0, res.stderr[-4000:]
    assert res.stdout.startswith("ok ")


def test_quad_rbgs_matches_direct_sweep(tmp_path):
    """The FP32 four-cells-per-lane colour kernel (csrc/rbgs_quad.h), run on
    the host over its launch grid, equals a cell-by-cell sweep bit for bit,
    with periodic axes, skipped solid faces, a pinned cell, zero starts and
    rows whose length is not a multiple of four; AddressSanitizer checks the
    16-byte loads past the row ends."""
    exe = str(tmp_path / "rbgs_quad_emul")
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + CSRC,
           os.path.join(HERE, "emul", "rbgs_quad_emul.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    res = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert res.returncode == 0, res.stderr[-4000:]
    assert res.stdout.startswith("ok ")
