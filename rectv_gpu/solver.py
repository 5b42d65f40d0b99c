"""The Python solver (the reference's src/rectv_gpu/solver.py)."""
from rectv_gpu_amd.solver import Solver, getp

__all__ = ["Solver", "getp"]
