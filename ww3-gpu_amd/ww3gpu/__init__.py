"""Python access to libww3gpu.so, the MI355X WAVEWATCH III spectral-step library.

The library's interface is its C ABI (include/ww3gpu_c.h), the same entry points the
reference's WW3 fork binds (kokkos/src/fortran_iface/ww_kokkos_c.hpp).  This package is
a thin ctypes layer over it for tests, benchmarks and Python callers; it computes
nothing itself.  There is no CPU fallback: if the library is missing, loading fails.
"""
from .capi import (  # noqa: F401
    LIB_PATH,
    Status,
    WW3GpuError,
    lib,
    load,
)
