"""Shared fixtures: the product library (HIP, C-ABI) and the CPU oracle."""
from __future__ import annotations

import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
PKG = importlib.import_module("2020researchproject_amd")
ORACLE_LIB = os.path.join(ROOT, "oracle", "libibm3d_oracle.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def I():
    return PKG.ibm3d


@pytest.fixture(scope="session")
def cases():
    return PKG.cases


@pytest.fixture(scope="session")
def oracle_lib(I):
    if not os.path.exists(ORACLE_LIB):
        pytest.fail("oracle not built: run __graft_entry__.build()")
    return I.load(ORACLE_LIB)


@pytest.fixture(scope="session")
def gpu_lib(I):
    return I.load()  # raises when the HIP library is missing


@pytest.fixture
def new_oracle(I, oracle_lib):
    made = []

    def make():
        s = I.Solver(oracle_lib)
        made.append(s)
        return s
    yield make
    for s in made:
        s.close()


@pytest.fixture
def new_gpu(I, gpu_lib):
    made = []

    def make():
        s = I.Solver(gpu_lib)
        made.append(s)
        return s
    yield make
    for s in made:
        s.close()
