"""MI355X-native drop-in for rectv_gpu (time-domain-decomposition 4D tomography).

``rectv`` and ``Solver`` mirror the reference package (src/rectv_gpu/__init__.py:3-4);
the numerics run in hand-written HIP kernels behind the C ABI of
include/rectv_amd.h (librectv_amd.so, built in-tree).
"""
from .rectv import rectv
from .solver import Solver, getp
from .basis import takephi, angles

__all__ = ["rectv", "Solver", "getp", "takephi", "angles"]
__version__ = "0.1.0"
