"""MI355X-native fractional-step ghost-cell IBM Navier-Stokes time step.

The product is the C-ABI shared library ``libibm3d_hip.so`` (HIP kernels for
gfx950 behind the reference's ``IBMSolver_*`` C API, see include/ibm3d.h).
``ibm3d`` is its ctypes binding; ``cases`` builds the benchmark workloads;
``serving`` drives the native request loop over streams of cases.
"""
from . import ibm3d, cases  # noqa: F401

__all__ = ["ibm3d", "cases", "serving"]
