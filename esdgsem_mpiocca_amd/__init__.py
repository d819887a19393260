"""MI355X-native shallow-water ESDGSEM hot path.

The per-stage residual and time step of the reference's deviceclass (ESDGSEM_MPIOCCA,
src/deviceclass.cpp) as hand-written gfx950 HIP kernels behind a C ABI (include/esdg.h), with the
host setup in C++ (include/esdg_setup.h).  Python is a thin ctypes layer for tests and tools.
"""
from .setup_api import Case, RunParams, basis, rk_coefficients, total_entropy, total_mass  # noqa: F401

__all__ = ["Case", "RunParams", "basis", "rk_coefficients", "total_entropy", "total_mass",
           "device"]


def device():
    """The HIP binding (imports libesdg.so; raises if it is missing)."""
    from . import device_api
    device_api.lib()
    return device_api
