"""The native solver class (the reference's SWIG module src/cuda/rectv.i)."""
from rectv_gpu_amd.rectv import rectv

__all__ = ["rectv"]
