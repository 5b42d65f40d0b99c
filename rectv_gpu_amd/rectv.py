"""``rectv``: the native solver class, with the method set of the reference's SWIG
class (src/cuda/rectv.i:9-28): constructor ``rectv(n, ntheta, m, nz, nzp, ngpus)``,
``run``, ``check_adjoints`` and ``free`` taking raw host addresses as integers,
and read-only ``n``, ``ntheta``, ``m``, ``nz``.
"""
import ctypes

from . import _lib


class rectv:
    """Proxy of the C-ABI solver handle (replaces rectv::rectv, src/cuda/rectv.cu:5-59)."""

    def __init__(self, n, ntheta, m, nz, nzp, ngpus):
        self._h = None
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().rtv_create(int(n), int(ntheta), int(m), int(nz), int(nzp), int(ngpus),
                                         ctypes.byref(h)))
        self._h = h
        self._n, self._ntheta, self._m, self._nz, self._nzp = int(n), int(ntheta), int(m), int(nz), int(nzp)

    # read-only attributes (rectv.i:13-19)
    n = property(lambda self: self._n)
    ntheta = property(lambda self: self._ntheta)
    m = property(lambda self: self._m)
    nz = property(lambda self: self._nz)
    nzp = property(lambda self: self._nzp)

    def _handle(self):
        if self._h is None:
            raise _lib.RectvError("solver has been freed")
        return self._h

    def run(self, fres, g, theta, phi, center, lambda0, lambda1, niter, titer):
        """rectv::run (src/cuda/rectv.cu:102-252); arguments are host addresses (ints)."""
        _lib.check(_lib.lib().rtv_run(self._handle(), fres, g, theta, phi, float(center), float(lambda0),
                                      float(lambda1), int(niter), int(titer)))

    def check_adjoints(self, res, g, theta, phi, lambda1, center):
        """rectv::check_adjoints (src/cuda/rectv.cu:255-289); res is the address of 6 doubles."""
        _lib.check(_lib.lib().rtv_check_adjoints(self._handle(), res, g, theta, phi, float(lambda1),
                                                 float(center)))

    def free(self):
        """rectv::free (src/cuda/rectv.cu:63-91); idempotent, never resets the device."""
        if self._h is not None:
            h, self._h = self._h, None
            _lib.check(_lib.lib().rtv_free(h))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
