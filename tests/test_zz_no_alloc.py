"""rtv_prox_noalloc (the closing ADMM phase on caller-owned device arrays, with its
scratch in a fixed per-device array) against rtv_apply_prox (the same phase on a
solver handle, with scratch the library allocates), through the C ABI: two calls
in a row, the second with smaller sizes, outputs equal to the handle-based entry
point's (the same phase: mu, rho and the history row bitwise; psi and div chi to
FP32 rounding, since the handle's solver stores chi = mu - rho psi in place of
psi and exports psi = (mu - chi) / rho); sizes beyond the scratch
are rejected before any launch.  Device buffers of the caller come from hipMalloc
through ctypes."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Dev:
    """A device copy of a host array (hipMalloc / hipMemcpy through ctypes)."""
    hip = None

    def __init__(self, host):
        if Dev.hip is None:
            Dev.hip = ctypes.CDLL("libamdhip64.so")
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        assert Dev.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.host.nbytes, 16))) == 0
        assert Dev.hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_size_t(self.host.nbytes), 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.host)
        assert Dev.hip.hipDeviceSynchronize() == 0
        assert Dev.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        Dev.hip.hipFree(self.ptr)


def _inputs(seed, n, m, nz, nzp):
    rng = np.random.default_rng(seed)
    nch = nz // nzp
    A = rng.standard_normal((nz, m, n, n)).astype(np.float32)
    fcur = (A + 0.3 * rng.standard_normal(A.shape)).astype(np.float32)
    shp = (nch, nzp + 1, m + 1, n + 1, n + 1, 4)
    psi = (0.2 * rng.standard_normal(shp)).astype(np.float32)
    mu = (0.1 * rng.standard_normal(shp)).astype(np.float32)
    return A, fcur, psi, mu


def _with_handle(n, m, nz, nzp, A, fcur, psi, mu, rho, lam0, lam1):
    from rectv_gpu_amd import Solver
    with Solver(n, 8, m, nz, nzp, 1) as s:
        s.set_state(A, psi, mu, rho, lam1)
        row = s.apply_prox(fcur, lam0, lam1)
        st = s.get_state()
    return row, st


def _noalloc(n, m, nz, nzp, A, fcur, psi, mu, rho, lam0, lam1):
    from rectv_gpu_amd import _lib
    bufs = dict(fcur=Dev(fcur), fprev=Dev(A), psi=Dev(psi), mu=Dev(mu), dchi=Dev(np.zeros_like(A)),
                rho=Dev(np.array([rho], np.float32)), row=Dev(np.zeros(7)))
    try:
        b = bufs
        _lib.check(_lib.lib().rtv_prox_noalloc(b["fcur"].ptr, b["fprev"].ptr, b["psi"].ptr, b["mu"].ptr,
                                               b["dchi"].ptr, b["rho"].ptr, b["row"].ptr, n, m, nz, nzp,
                                               float(lam0), float(lam1), None))
        return {k: v.get() for k, v in bufs.items()}
    finally:
        for v in bufs.values():
            v.free()


def test_noalloc_matches_handle_entry_point_twice(gpu):
    # the second call is smaller than the first: the shared scratch keeps the first
    # call's larger contents, which must not leak into the second's result
    for k, (n, m, nz, nzp, rho, lam1) in enumerate([(64, 5, 6, 2, 0.75, 2.5), (32, 3, 2, 1, 2.0, 4.0)]):
        say("case rtv_prox_noalloc call", k + 1, "n m nz nzp", n, m, nz, nzp)
        A, fcur, psi, mu = _inputs(100 + k, n, m, nz, nzp)
        lam0 = 0.3 * rho
        row_h, st_h = _with_handle(n, m, nz, nzp, A, fcur, psi, mu, rho, lam0, lam1)
        out = _noalloc(n, m, nz, nzp, A, fcur, psi, mu, rho, lam0, lam1)
        np.testing.assert_array_equal(out["mu"], st_h["mu"], err_msg="mu")
        # the handle's solver keeps chi = mu - rho psi in place of psi: its psi comes back
        # as (mu - chi) / rho and its div chi from chi, equal to FP32 rounding
        for key in ("psi", "dchi"):
            e = float(np.linalg.norm(out[key].astype(np.float64) - st_h[key]) / np.linalg.norm(st_h[key]))
            say("  %s rel L2 %.2e" % (key, e))
            assert e < 1e-6, key
        np.testing.assert_array_equal(out["rho"][0], np.float32(st_h["rho"]))
        np.testing.assert_array_equal(out["row"], row_h)  # NaN data term in both
        say("  equal: mu, rho, row", row_h)


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    from rectv_gpu_amd import _lib
    L = _lib.lib()
    one = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    # n=2048, m=32, nz=64, nzp=16: 129*33*17*4 prox workgroups of 5 sums = 11.6 MB > 8 MiB
    st = L.rtv_prox_noalloc(one, one, one, one, one, one, one, 2048, 32, 64, 16, 1e-3, 4.0, None)
    assert st == -1 and b"scratch" in L.rtv_last_error()
