"""ctypes binding of librectv_amd.so (the C ABI declared in include/rectv_amd.h).

The library is built in-tree by ``python -m rectv_gpu_amd.build``.  There is no
fallback: if the shared library is missing or cannot be loaded, importing the
solver classes raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# Batched and served jobs run on their own streams; HIP maps a process's streams onto
# GPU_MAX_HW_QUEUES hardware queues (4 by default), which caps how many jobs' kernels
# overlap (16 queues shorten the 64-job mix by about 4%).  Set before the library and
# the HIP runtime load; a value the caller set is kept.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
# RECTV_LIB selects another build of the same library (e.g. one compiled with other
# build-time tuning constants); by default the in-tree build is loaded.
LIB_PATH = os.environ.get("RECTV_LIB") or os.path.join(_HERE, "librectv_amd.so")

_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float

# name -> (restype, argtypes); must match include/rectv_amd.h
SIGNATURES = {
    "rtv_last_error": (ctypes.c_char_p, []),
    "rtv_device_count": (_I, []),
    "rtv_create": (_I, [_I, _I, _I, _I, _I, _I, ctypes.POINTER(_P)]),
    "rtv_free": (_I, [_P]),
    "rtv_run": (_I, [_P, _P, _P, _P, _P, _F, _F, _F, _I, _I]),
    "rtv_check_adjoints": (_I, [_P, _P, _P, _P, _P, _F, _F]),
    "rtv_set_inputs": (_I, [_P, _P, _P, _P, _F]),
    "rtv_reset": (_I, [_P]),
    "rtv_iterate": (_I, [_P, _I, _I, _F, _F, _P]),
    "rtv_get_result": (_I, [_P, _P]),
    "rtv_synchronize": (_I, [_P]),
    "rtv_set_state": (_I, [_P, _P, _P, _P, _F, _F]),
    "rtv_get_state": (_I, [_P, _P, _P, _P, _P, _P]),
    "rtv_apply_inner_step": (_I, [_P, _P, _P, _F]),
    "rtv_apply_prox": (_I, [_P, _P, _F, _F, _P]),
    "rtv_prox_noalloc": (_I, [_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _F, _F, _P]),
    "rtv_set_support": (_I, [_P, _F]),
    "rtv_apply_fbp": (_I, [_P, _P, _P]),
    "rtv_fbp_start": (_I, [_P]),
    "rtv_apply_radon": (_I, [_P, _P, _P]),
    "rtv_apply_radon_adj": (_I, [_P, _P, _P]),
    "rtv_apply_gradient": (_I, [_P, _P, _I, _F, _P]),
    "rtv_apply_divergence": (_I, [_P, _P, _F, _F, _P]),
    "rtv_set_profiling": (_I, [_P, _I]),
    "rtv_get_profile": (_I, [_P, ctypes.c_char_p, _I]),
    "rtv_device_bytes": (ctypes.c_longlong, [_P]),
    "rtv_iterate_async": (_I, [_P, _I, _I, _F, _F]),
    "rtv_query": (_I, [_P, ctypes.POINTER(_I)]),
    "rtv_iterations": (_I, [_P, ctypes.POINTER(ctypes.c_longlong)]),
    "rtv_history": (_I, [_P, ctypes.c_longlong, _I, _P]),
    "rtv_set_graph": (_I, [_P, _I]),
    "rtv_autotune": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "rtv_batch_iterate": (_I, [ctypes.POINTER(_P), _I, _P, _P, _P, _P]),
    "rtv_footprint": (_I, [_I, _I, _I, _I, _I, ctypes.POINTER(ctypes.c_longlong)]),
    "rtv_prep_darkflat": (_I, [_P, _P, _P, _I, _I, _I, _I, _I, _I, _P]),
    "rtv_prep_stripe_ti": (_I, [_P, _I, _I, _I, _F, _I]),
    "rtv_prep_paganin": (_I, [_P, _I, _I, _I, _F, _F, _F, _F]),
}


class RectvError(RuntimeError):
    """A C-ABI call returned a failure status."""


def lib():
    """Load (once) and return the shared library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("librectv_amd.so is not built; run `python -m rectv_gpu_amd.build` "
                              "(expected at %s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        msg = lib().rtv_last_error().decode("utf-8", "replace")
        if status == -1:
            raise ValueError(msg)
        raise RectvError("status %d: %s" % (status, msg))
    return status


def device_count():
    return lib().rtv_device_count()
