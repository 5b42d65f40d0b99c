"""Drop-in import name of the reference package (src/rectv_gpu/__init__.py:3-4).

``import rectv_gpu`` gives the MI355X-native classes of ``rectv_gpu_amd``: as in
the reference, ``rectv_gpu.rectv`` is the native solver class and
``rectv_gpu.Solver`` the Python solver, so code written against the reference
(e.g. its tests/test_adjoint.py) runs unchanged.
"""
from rectv_gpu.rectv import *  # noqa: F401,F403  (rebinds rectv_gpu.rectv to the class)
from rectv_gpu.solver import *  # noqa: F401,F403
