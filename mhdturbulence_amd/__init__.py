"""MI355X-native MHD time step (MHDTurbulence's src_cpp_kokkos_host step, drop-in).

The compute path is libmhd_amd.so (hand-written HIP kernels for gfx950 behind the C ABI
in include/mhd_amd.h); this package binds it and mirrors the reference interface.
"""
from .solver import Batch, Box, Config, PERIODIC, REFLECTION, OUTFLOW  # noqa: F401

__all__ = ["Batch", "Box", "Config", "PERIODIC", "REFLECTION", "OUTFLOW"]
