"""MI355X-native ray casting for RobartsVTK's CUDA volume mappers.

The compute path is librvkhip.so (HIP kernels for gfx950 behind the C-ABI in
include/rvk.h); `abi` is its ctypes view and `configs` builds the seeded
benchmark workloads.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
