"""Python drop-in for rectv_gpu.Solver (src/rectv_gpu/solver.py:11-93).

Same constructor, ``recon``/``adjoint_tests`` signatures, defaults and return
arrays as the reference; inputs are validated (dtype float32, C order, shapes)
and converted where that is lossless, where the reference would read whatever
bytes the pointer gives.  Extra methods (``set_inputs``, ``iterate``, ...) expose
the staged form used for timing.
"""
import ctypes

import numpy as np

from . import _lib
from .rectv import rectv


def getp(a):
    """Host address of a numpy array (src/rectv_gpu/solver.py:7-8)."""
    return a.__array_interface__['data'][0]


def _f32(a, shape, name):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != tuple(shape):
        if a.size == int(np.prod(shape)):
            a = a.reshape(shape)
        else:
            raise ValueError("%s has shape %s, expected %s" % (name, a.shape, tuple(shape)))
    return a


class Solver(rectv):
    """Time-resolved tomography solver (ADMM with 4D total variation).

    Parameters are those of the reference: n (detector width / slice size),
    ntheta (total number of projections), m (number of temporal basis
    functions), nz (slices), nzp (slices per chunk), ngpus (accepted; this library
    drives one GPU).
    """

    def __init__(self, n, ntheta, m, nz, nzp, ngpus):
        super().__init__(n, ntheta, m, nz, nzp, ngpus)

    def __enter__(self):
        return self

    def __exit__(self, type, value, traceback):
        self.free()

    def _inputs(self, data, theta, phi):
        data = _f32(data, (self.nz, self.ntheta, self.n), "data")
        theta = _f32(theta, (self.ntheta,), "theta")
        phi = _f32(phi, (self.m, 2 * self.ntheta), "phi")
        return data, theta, phi

    def recon(self, data, theta, phi, rot_center=None, lambda0=1e-3, lambda1=1, niter=1, titer=1, *,
              init=None, support=None):
        """Reconstruct (src/rectv_gpu/solver.py:45-84).

        data [nz, ntheta, n], theta [ntheta], phi [m, 2*ntheta] (re/im
        interleaved); returns float32 [nz, m, n, n].  As in the reference the
        TV weight passed to the native solver is lambda0*lambda1 and the
        rotation centre defaults to n/2.

        Keyword-only extensions (defaults give the reference's iteration):
        init="fbp" starts from the filtered back-projection of the data
        (rtv_fbp_start) instead of zero; support=r >= 0 keeps every frame inside
        the circ(r) mask after each inner step (rtv_set_support).
        """
        lambda0 = lambda0 * lambda1
        if rot_center is None:
            rot_center = self.n / 2
        if init not in (None, "zero", "fbp"):
            raise ValueError("init must be None, 'zero' or 'fbp'")
        data, theta, phi = self._inputs(data, theta, phi)
        rtv = np.zeros([self.nz, self.m, self.n, self.n], dtype='float32')
        if init in (None, "zero") and support is None:
            self.run(getp(rtv), getp(data), getp(theta), getp(phi), rot_center, lambda0, lambda1, niter, titer)
            return rtv
        self.set_support(support)
        try:
            self.set_inputs(data, theta, phi, rot_center)
            if init == "fbp":
                _lib.check(_lib.lib().rtv_fbp_start(self._handle()))
            else:
                self.reset()
            self.iterate(niter, titer, lambda0, lambda1)
            return self.result()
        finally:
            self.set_support(None)

    def set_support(self, r):
        """Support mask circ(r) (kernels_radonusfft.cuh:146-159) on the filtered
        back-projection and after every inner step; None or r < 0 removes it."""
        _lib.check(_lib.lib().rtv_set_support(self._handle(), -1.0 if r is None else float(r)))

    def fbp(self, data, theta=None, phi=None, rot_center=None):
        """Filtered back-projection of data [nz, ntheta, n] as a static object, the same
        image in each of the m frames: pi/(4 sqrt(m)) Re adjR_w(data) with the applyfilter
        weights (kernels_radonusfft.cuh:173-189), masked by the support if set.  theta and
        phi (with rot_center) set the geometry first when given."""
        if theta is not None:
            self.set_inputs(data, theta, phi, rot_center)
        g = _f32(data, (self.nz, self.ntheta, self.n), "data")
        f = np.zeros((self.nz, self.m, self.n, self.n), np.float32)
        _lib.check(_lib.lib().rtv_apply_fbp(self._handle(), getp(g), getp(f)))
        return f

    def fbp_start(self):
        """Start the resident state from the filtered back-projection of the resident data."""
        _lib.check(_lib.lib().rtv_fbp_start(self._handle()))

    def adjoint_tests(self, data, theta, phi, lambda1, rot_center=None):
        """The six inner products of rectv::check_adjoints (src/rectv_gpu/solver.py:86-93)."""
        if rot_center is None:
            rot_center = self.n / 2
        data, theta, phi = self._inputs(data, theta, phi)
        res = np.zeros([6])
        self.check_adjoints(getp(res), getp(data), getp(theta), getp(phi), lambda1, rot_center)
        return res

    # ---- staged API (inputs resident on the device) ---------------------------
    def set_inputs(self, data, theta, phi, rot_center=None):
        if rot_center is None:
            rot_center = self.n / 2
        data, theta, phi = self._inputs(data, theta, phi)
        _lib.check(_lib.lib().rtv_set_inputs(self._handle(), getp(data), getp(theta), getp(phi),
                                             float(rot_center)))

    def reset(self):
        _lib.check(_lib.lib().rtv_reset(self._handle()))

    def iterate(self, niter, titer, lambda0, lambda1, history=False):
        """Run niter outer iterations on the resident state (lambda0 used as given).
        With history=True returns an array [niter, 4] of (r, s, rho, ||f(k+1)-f(k)||^2)."""
        hist = np.zeros((max(niter, 1), 4)) if history else None
        _lib.check(_lib.lib().rtv_iterate(self._handle(), int(niter), int(titer), float(lambda0),
                                          float(lambda1), getp(hist) if history else None))
        return hist[:niter] if history else None

    def iterate_async(self, niter, titer, lambda0, lambda1):
        """Enqueue niter outer iterations on the solver's stream and return at once."""
        _lib.check(_lib.lib().rtv_iterate_async(self._handle(), int(niter), int(titer), float(lambda0),
                                                float(lambda1)))

    def idle(self):
        """True when every enqueued iteration has finished."""
        v = ctypes.c_int()
        _lib.check(_lib.lib().rtv_query(self._handle(), ctypes.byref(v)))
        return bool(v.value)

    def iterations(self):
        """Outer iterations enqueued since the last reset."""
        v = ctypes.c_longlong()
        _lib.check(_lib.lib().rtv_iterations(self._handle(), ctypes.byref(v)))
        return int(v.value)

    HIST_COLUMNS = ("r", "s", "rho", "rdiff", "fnorm2", "data", "tv")

    def history(self, first, count):
        """Rows [first, first+count) since reset as [count, 7]: r, s, rho used,
        ||f(k+1)-f(k)||^2, ||f(k+1)||^2, 1/2||R f(k) - g||^2 of the entering iterate
        (NaN when titer = 0) and sum |grad f(k+1)| (waits for the solver's stream)."""
        out = np.zeros((max(count, 1), len(self.HIST_COLUMNS)))
        _lib.check(_lib.lib().rtv_history(self._handle(), int(first), int(count), getp(out)))
        return out[:count]

    @staticmethod
    def objective(rows, lambda0):
        """Objective J(f_k) = 1/2||R f_k - g||^2 + lambda0/2 sum|grad f_k| of the iterates
        f_1 .. f_{count-1} from history rows 0 .. count-1 (lambda0 as given to
        iterate, i.e. already multiplied by lambda1)."""
        rows = np.asarray(rows)
        return rows[1:, 5] + 0.5 * lambda0 * rows[:-1, 6]

    # ---- state (checkpoint / restore) -------------------------------------------
    def dual_shape(self):
        return (self.nz // self.nzp, self.nzp + 1, self.m + 1, self.n + 1, self.n + 1, 4)

    def set_state(self, f=None, psi=None, mu=None, rho=0.5, lambda1=1.0):
        """Resume from a saved state (None: zero); see rtv_set_state."""
        V = (self.nz, self.m, self.n, self.n)
        f = None if f is None else _f32(f, V, "f")
        psi = None if psi is None else _f32(psi, self.dual_shape(), "psi")
        mu = None if mu is None else _f32(mu, self.dual_shape(), "mu")
        p = lambda a: None if a is None else getp(a)
        _lib.check(_lib.lib().rtv_set_state(self._handle(), p(f), p(psi), p(mu), float(rho), float(lambda1)))

    def get_state(self):
        """dict(f, psi, mu, dchi, rho) of the resident state."""
        V = (self.nz, self.m, self.n, self.n)
        out = dict(f=np.zeros(V, np.float32), psi=np.zeros(self.dual_shape(), np.float32),
                   mu=np.zeros(self.dual_shape(), np.float32), dchi=np.zeros(V, np.float32))
        rho = np.zeros(1, np.float32)
        _lib.check(_lib.lib().rtv_get_state(self._handle(), getp(out["f"]), getp(out["psi"]), getp(out["mu"]),
                                            getp(out["dchi"]), getp(rho)))
        out["rho"] = float(rho[0])
        return out

    def apply_inner_step(self, fcur, lambda1):
        V = (self.nz, self.m, self.n, self.n)
        fcur = _f32(fcur, V, "fcur")
        out = np.zeros(V, np.float32)
        _lib.check(_lib.lib().rtv_apply_inner_step(self._handle(), getp(fcur), getp(out), float(lambda1)))
        return out

    def apply_prox(self, fcur, lambda0, lambda1):
        """Returns the history row (7 values) of the closing phase applied to fcur."""
        fcur = _f32(fcur, (self.nz, self.m, self.n, self.n), "fcur")
        row = np.zeros(len(self.HIST_COLUMNS))
        _lib.check(_lib.lib().rtv_apply_prox(self._handle(), getp(fcur), float(lambda0), float(lambda1), getp(row)))
        return row

    def set_graph(self, on=True):
        """Capture each outer iteration once as a HIP graph and replay it."""
        _lib.check(_lib.lib().rtv_set_graph(self._handle(), 1 if on else 0))

    def autotune(self):
        """Pick the gather/scatter launch shapes for this geometry by timing them;
        returns (gather slices per workgroup, scatter slices per workgroup)."""
        a, b = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().rtv_autotune(self._handle(), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    @staticmethod
    def footprint(n, ntheta, m, nz, nzp):
        """Upper bound of the device bytes a solver of these sizes holds."""
        v = ctypes.c_longlong()
        _lib.check(_lib.lib().rtv_footprint(int(n), int(ntheta), int(m), int(nz), int(nzp), ctypes.byref(v)))
        return int(v.value)

    def result(self):
        out = np.zeros([self.nz, self.m, self.n, self.n], dtype='float32')
        _lib.check(_lib.lib().rtv_get_result(self._handle(), getp(out)))
        return out

    def synchronize(self):
        _lib.check(_lib.lib().rtv_synchronize(self._handle()))

    # ---- operator entry points (parity tests) ----------------------------------
    def apply_radon(self, f):
        f = _f32(f, (self.nz, self.m, self.n, self.n), "f")
        g = np.zeros((self.nz, self.ntheta, self.n), np.float32)
        _lib.check(_lib.lib().rtv_apply_radon(self._handle(), getp(f), getp(g)))
        return g

    def apply_radon_adj(self, g):
        g = _f32(g, (self.nz, self.ntheta, self.n), "g")
        f = np.zeros((self.nz, self.m, self.n, self.n), np.float32)
        _lib.check(_lib.lib().rtv_apply_radon_adj(self._handle(), getp(g), getp(f)))
        return f

    def apply_gradient(self, fext, iz, lambda1):
        fext = _f32(fext, (self.nzp + 2, self.m, self.n, self.n), "fext")
        h2 = np.zeros((self.nzp + 1, self.m + 1, self.n + 1, self.n + 1, 4), np.float32)
        _lib.check(_lib.lib().rtv_apply_gradient(self._handle(), getp(fext), int(iz), float(lambda1), getp(h2)))
        return h2

    def apply_divergence(self, h2, lambda1, tau, f):
        h2 = _f32(h2, (self.nzp + 1, self.m + 1, self.n + 1, self.n + 1, 4), "h2")
        f = _f32(f, (self.nzp, self.m, self.n, self.n), "f").copy()
        _lib.check(_lib.lib().rtv_apply_divergence(self._handle(), getp(h2), float(lambda1), float(tau), getp(f)))
        return f

    def set_profiling(self, on=True):
        _lib.check(_lib.lib().rtv_set_profiling(self._handle(), 1 if on else 0))

    def profile(self):
        """{kernel: (launches, total_ms)} from HIP events on the solver's stream."""
        buf = ctypes.create_string_buffer(1 << 16)
        _lib.check(_lib.lib().rtv_get_profile(self._handle(), buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def device_bytes(self):
        return int(_lib.lib().rtv_device_bytes(self._handle()))
