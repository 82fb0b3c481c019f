"""MI355X-native finite-volume step for the 2-D compressible Euler equations.

A drop-in for the FP64 CPU/OpenMP step of RP_FVM (xujf520/OverheatingProblem):
hand-written HIP kernels for gfx950 behind the C ABI in include/rpfvm.h.
"""
from .config import Config, make_config  # noqa: F401

__version__ = "0.1.0"
