"""MI355X-native WRAPD: the ADMM elasticity solver on hand-written HIP kernels.

Importing the package does not touch the GPU.  `wrapd_amd.solver.Solver` drives the
C++ host classes and the device kernels of wrapd_amd/lib/libwrapd.so.
"""
from . import meshgen  # noqa: F401

__all__ = ["meshgen", "solver"]
