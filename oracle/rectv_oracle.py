"""CPU restatement of rectv_gpu's 4D reconstruction path (NumPy, FP64 by default).

TEST INFRASTRUCTURE ONLY.  The HIP product path is checked against this module:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product (``rectv_gpu_amd``) never imports it and
has no CPU fallback.

Every function restates one piece of the reference, step by step, in the order
the reference performs it (file:line citations are relative to the reference
repository).  The restatement keeps the reference's own data shapes, including
the (2n+2K)^2 padded grid with its periodic border, the checkerboard
multiplications, the per-basis-function loops and the ghost/halo dual entries,
so that the restructured GPU pipeline is checked against the literal algorithm.

Pinning (what makes this oracle trustworthy, see tests/test_oracle_pins.py):
  * 1D/2D DFT conventions: checked against explicit DFT matrices built from the
    definition (the brute-force NUDFT test of ``test_oracle_pins.py``).
  * USFFT gridding (deapodisation, oversampled FFT, wrap, Gaussian gather):
    checked against a brute-force nonuniform DFT at the polar points, at the
    USFFT accuracy eps=1e-3, and against analytic line integrals of a Gaussian
    blob (orientation and scale of the projection geometry).
  * adjR / radonapradj: adjointness <R u, v> = <u, R* v> to FP64 rounding.
  * decphi with the Fourier basis: equals a centred DFT across time frames.
  * M=1, constant basis: radonapr reduces to Re(fwdR).
  * gradf: equals forward differences with clamped (Neumann) borders computed
    with np.diff; div: the summation-by-parts identity <f, div h> = 1/2 <grad f, h>.
  * solve_reg: the group soft-threshold is the minimiser of 1/2|p-v|^2 + t|p|
    (checked by numerical minimisation).
  * ADMM: invariants (lambda0=0 keeps mu at 0, niter=0 returns zeros), descent of
    the data residual on a noise-free problem.
  * check_adjoints: res[0] == res[1] to FP64 rounding.
  * solve level (tests/test_oracle_solve_pins.py): basis orientation (with M =
    ntheta the takephi basis is a full DFT, so row k of R_apr f is sqrt(M) Re
    fwdR(f_k) row k), frame times t_i against hand-evaluated C integer division,
    chunk-order invariance of run, the rho rule's direction against residuals
    computed from the iterates (residual balancing), and the fixed-point
    conditions of a converged run (psi = grad f, mu in the group subdifferential,
    R^*(Rf - g) + div mu = 0).  tests/test_oracle_mutations.py plants five
    plausible mistakes (time-reversed basis, inverted rho rule, dual-update sign,
    halo from the new iterate, rounded frame times) in copies of this module and
    checks that each fails one of these pins.
  * FBP filter, support mask, fbp (tests/test_oracle_fbp.py): a hand-counted
    mask, the filter's symmetry and ramp limit, the filter equal to an explicit
    circular convolution on the detector before adjR, an analytic disk
    reconstructed with contrast 1, the warm start lowering the objective.
  * pre-processing (tests/test_oracle_prep.py): dark/flat -log inverting a known
    attenuation; the stripe removal's closed-form response to a single stripe;
    the phase filter's unit gain at zero frequency and its large-alpha limit.
    retrieve_phase restates tomopy's published retrieve_phase (tomopy.prep.phase,
    called through tomocupy's command line of examples/tao/rec_line.txt:1; no
    version is pinned by the reference and neither package is installed here):
    its padding widths, reciprocal grid and max normalisation are parity
    unpinned beyond the properties above.
Every other function here has at least one of these checks.

Precision: FP64 throughout by default (``dtype=np.float64``).  With
``dtype=np.float32`` the same code runs in FP32/complex64; that variant is the
CPU baseline timed by ``bench.py`` (``kind: "port"``).
"""
from __future__ import annotations

import sys
import numpy as np
import scipy.fft as sfft
import scipy.sparse as sp

PI_REF = 3.1415926535  # kernels_radonusfft.cuh:1 (the USFFT's truncated PI)

_WORKERS = None  # scipy.fft worker threads (None = scipy default, 1)


def set_workers(w):
    """Number of threads scipy.fft uses (CPU baseline timing uses all cores)."""
    global _WORKERS
    _WORKERS = w


def _cdt(dtype):
    return np.complex128 if np.dtype(dtype) == np.float64 else np.complex64


# ---------------------------------------------------------------------------
# radonusfft  (radonusfft.cu / kernels_radonusfft.cuh)
# ---------------------------------------------------------------------------

def usfft_params(n):
    """USFFT constants, radonusfft.cu:10-12.

    eps = 1e-3 (float); mu = -log(eps)/(2 n^2) stored as float; the kernel
    half-width K = ceil(2n/PI * sqrt(-mu log eps + (mu n)^2/4)) (= 4 for every n).
    """
    eps = np.float32(1e-3)
    mu = np.float32(-np.log(np.float64(eps)) / (2.0 * n * n))
    mu64 = float(mu)
    K = int(np.ceil(2 * n * 1 / PI_REF * np.sqrt(-mu64 * np.log(np.float64(eps))
                                                  + (mu64 * n) * (mu64 * n) / 4)))
    return mu64, K


def sigma(idx):
    """Checkerboard factor 1 - 2*((i+1) % 2) used by fftshift, kernels_radonusfft.cuh:33-35."""
    idx = np.asarray(idx)
    return (1 - 2 * ((idx + 1) % 2)).astype(np.float64)


def takexy(theta, n):
    """Polar sample coordinates, kernels_radonusfft.cuh:66-79.

    x[k, s] = (s - n/2)/n * sin(theta_k), y = (s - n/2)/n * cos(theta_k);
    values >= 0.5 are clamped to 0.5 - 1e-5.  theta arrives as float32 and is
    used here exactly (converted to FP64).
    """
    th = np.asarray(theta, dtype=np.float32).astype(np.float64)
    s = (np.arange(n) - n // 2) / n
    x = s[None, :] * np.sin(th)[:, None]
    y = s[None, :] * np.cos(th)[:, None]
    x = np.where(x >= 0.5, 0.5 - 1e-5, x)
    y = np.where(y >= 0.5, 0.5 - 1e-5, y)
    return x, y


def takeshift(c, n):
    """Centre-shift phase table exp(i 2 PI c (s - n/2)/n), kernels_radonusfft.cuh:191-198.

    Called with c = -(center - n/2) for the forward and +(center - n/2) for the
    adjoint (radonusfft.cu:70-75).
    """
    s = np.arange(n)
    a = 2 * PI_REF * c * (s - n / 2.0) / n
    return np.cos(a) + 1j * np.sin(a)


class Usfft:
    """The radonusfft object for one (n, ntheta) geometry (radonusfft.cu:5-55).

    The Gaussian gather of kernels_radonusfft.cuh:81-144 is held as a sparse
    matrix over the padded (2n+2K)^2 grid (81 taps per polar sample), so that the
    forward gather is W @ grid and the adjoint scatter is W^T @ samples; the
    taps, weights and the s == 0 zeroing are exactly those of the kernel.
    """

    def __init__(self, n, theta, center, dtype=np.float64):
        self.n = n
        self.ntheta = len(theta)
        self.dtype = np.dtype(dtype)
        self.cdtype = _cdt(dtype)
        self.mu, self.K = usfft_params(n)
        self.Ng = 2 * n + 2 * self.K
        self.x, self.y = takexy(theta, n)
        c = np.float32(center - n / 2.0)
        self.shiftfwd = takeshift(-float(c), n).astype(self.cdtype)
        self.shiftadj = takeshift(float(c), n).astype(self.cdtype)
        self._build_gather()
        # deapodisation factor phi (kernels_radonusfft.cuh:11)
        t = np.arange(n) - n // 2
        self.phi = np.exp(-self.mu * t[None, :] ** 2 - self.mu * t[:, None] ** 2)

    def _build_gather(self):
        n, K, Ng, mu = self.n, self.K, self.Ng, self.mu
        x0 = self.x.reshape(-1)
        y0 = self.y.reshape(-1)
        i = np.arange(2 * K + 1)
        ell0 = np.floor(2 * n * x0)[:, None] - K + i[None, :]      # fast / x axis
        ell1 = np.floor(2 * n * y0)[:, None] - K + i[None, :]      # slow / y axis
        w0 = ell0 / (2 * n) - x0[:, None]
        w1 = ell1 / (2 * n) - y0[:, None]
        # w = PI/mu * exp(-PI^2/mu * (w0^2 + w1^2))   kernels_radonusfft.cuh:118
        W = (PI_REF / mu) * np.exp(-PI_REF * PI_REF / mu *
                                   (w0[:, None, :] ** 2 + w1[:, :, None] ** 2))
        f_ind = (n + K + ell0[:, None, :]) + Ng * (n + K + ell1[:, :, None])
        ns = x0.size
        s_idx = np.arange(ns) % n
        W[s_idx == 0] = 0.0                       # s == 0 sample is zeroed (:104-108, :138-142)
        rows = np.repeat(np.arange(ns), (2 * K + 1) ** 2)
        self.W = sp.csr_matrix((W.reshape(-1), (rows, f_ind.reshape(-1).astype(np.int64))),
                               shape=(ns, Ng * Ng)).astype(self.dtype)
        self.WT = self.W.T.tocsr()

    # -- stages -------------------------------------------------------------
    def _sig2(self):
        a = sigma(np.arange(self.Ng))
        return a[:, None] * a[None, :]

    def fwdR(self, u):
        """Forward USFFT Radon, radonusfft.cu:77-104.  u [B, n, n] complex -> [B, ntheta, n]."""
        n, K, Ng = self.n, self.K, self.Ng
        B = u.shape[0]
        fde = np.zeros((B, Ng, Ng), dtype=self.cdtype)
        # divphi FWD (kernels_radonusfft.cuh:11-18)
        fde[:, n // 2 + K:n // 2 + K + n, n // 2 + K:n // 2 + K + n] = u / self.phi / (4 * n * n)
        s2 = self._sig2()
        fde *= s2                                                   # fftshift 2D (:87)
        sub = fde[:, K:K + 2 * n, K:K + 2 * n]
        fde[:, K:K + 2 * n, K:K + 2 * n] = sfft.fft2(sub, workers=_WORKERS)   # CUFFT_FORWARD (:89)
        fde *= s2                                                   # fftshift 2D (:90)
        fde = self._wrap_fwd(fde)                                   # wrap FWD (:92)
        g = (self.W @ fde.reshape(B, Ng * Ng).T).T.reshape(B, self.ntheta, n)  # gather FWD (:93)
        g = g / n
        g *= self.shiftfwd[None, None, :]                            # shift (:96)
        sg = sigma(np.arange(n))
        g *= sg                                                     # fftshift 1D (:97)
        g = sfft.ifft(g, axis=-1, workers=_WORKERS) * n             # CUFFT_INVERSE, unnormalised (:99)
        g *= sg                                                     # fftshift 1D (:100)
        g *= 1.0 / np.sqrt(n * self.ntheta)                         # mulr (:102)
        return g.astype(self.cdtype)

    def adjR(self, g, filter=False):
        """Adjoint USFFT Radon, radonusfft.cu:106-135.  g [B, ntheta, n] -> [B, n, n].

        filter=True multiplies the centred detector spectrum by the applyfilter
        weights (kernels_radonusfft.cuh:173-189) right after the forward FFT and its
        checkerboard: the `filter` argument of radonusfft::adjR (radonusfft.cu:106,
        radonusfft.cuh:37), which the reference's body leaves unused."""
        n, K, Ng = self.n, self.K, self.Ng
        B = g.shape[0]
        sg = sigma(np.arange(n))
        g = g * sg                                                  # fftshift 1D (:115)
        g = sfft.fft(g, axis=-1, workers=_WORKERS)                  # CUFFT_FORWARD (:117)
        g *= sg                                                     # fftshift 1D (:118)
        if filter:
            g = g * applyfilter_weights(n)[None, None, :]
        g *= self.shiftadj[None, None, :]                            # shift (:121)
        g = g / n                                                   # gather ADJ g0 = g/n (:102-103)
        fde = (self.WT @ g.reshape(B, self.ntheta * n).T).T.reshape(B, Ng, Ng)  # scatter (:126-131)
        fde = fde.astype(self.cdtype)
        fde = self._wrap_adj(fde)                                   # wrap ADJ (:123)
        s2 = self._sig2()
        fde *= s2                                                   # fftshift 2D (:125)
        sub = fde[:, K:K + 2 * n, K:K + 2 * n]
        fde[:, K:K + 2 * n, K:K + 2 * n] = sfft.ifft2(sub, workers=_WORKERS) * (4 * n * n)  # CUFFT_INVERSE (:128)
        fde *= s2                                                   # fftshift 2D (:129)
        f = fde[:, n // 2 + K:n // 2 + K + n, n // 2 + K:n // 2 + K + n] / self.phi / (4 * n * n)  # divphi ADJ
        f = f * (1.0 / np.sqrt(n * self.ntheta))                    # mulr (:132)
        return f.astype(self.cdtype)

    def _border_maps(self):
        n, K, Ng = self.n, self.K, self.Ng
        t = np.arange(Ng)
        src = (t - K + 2 * n) % (2 * n) + K
        border = (t < K) | (t >= 2 * n + K)
        return src, border

    def _wrap_fwd(self, fde):
        """Periodic border copy, kernels_radonusfft.cuh:40-58 (TOMO_FWD)."""
        src, border = self._border_maps()
        cell_border = border[:, None] | border[None, :]
        yy, xx = np.nonzero(cell_border)
        fde[:, yy, xx] = fde[:, src[yy], src[xx]]
        return fde

    def _wrap_adj(self, fde):
        """Fold the border into its periodic image, kernels_radonusfft.cuh:59-62 (TOMO_ADJ)."""
        src, border = self._border_maps()
        cell_border = border[:, None] | border[None, :]
        yy, xx = np.nonzero(cell_border)
        B = fde.shape[0]
        flat = fde.reshape(B, -1)
        np.add.at(flat, (slice(None), src[yy] * self.Ng + src[xx]), flat[:, yy * self.Ng + xx])
        return flat.reshape(fde.shape)


def applyfilter_weights(n):
    """FBP filter of kernels_radonusfft.cuh:173-189 on the centred detector spectrum:
    rho = (s - n/2)/n (integer n/2), w = 0 at rho == 0, else |rho| n 4 sin(rho)/rho."""
    rho = (np.arange(n) - n // 2) / float(n)
    safe = np.where(rho == 0, 1.0, rho)
    return np.where(rho == 0, 0.0, np.abs(rho) * n * 4 * np.sin(safe) / safe)


def circ_mask(n, r):
    """Support mask of kernels_radonusfft.cuh:146-159, evaluated in float32 as there:
    x = (tx - n/2)/float(n), y likewise, kept where 4x^2 + 4y^2 < 1 - r."""
    t = ((np.arange(n) - n // 2).astype(np.float32) / np.float32(n)).astype(np.float32)
    q = np.float32(4) * t[None, :] * t[None, :] + np.float32(4) * t[:, None] * t[:, None]
    return (q < np.float32(1) - np.float32(r)).astype(np.float64)


def circ(f, r):
    """f[..., y, x] *= circ mask (kernels_radonusfft.cuh:146-159)."""
    return f * circ_mask(f.shape[-1], r)


# ---------------------------------------------------------------------------
# rectv operators (operators.cu / kernels_operators.cuh)
# ---------------------------------------------------------------------------

def phi_complex(phi):
    """The basis as complex [m, ntheta] from the float32 [m, 2*ntheta] re/im layout (rectv.cu:116)."""
    phi = np.asarray(phi, dtype=np.float32).astype(np.float64)
    return phi[:, 0::2] + 1j * phi[:, 1::2]


def tframes(ntheta, m):
    """Projection index of time frame i: i*ntheta/m in C int division (kernels_operators.cuh:113)."""
    return (np.arange(m) * ntheta) // m


def decphi(fc, phij, ntheta, m):
    """u = sum_i f_i conj(phi_j[t_i]), kernels_operators.cuh:99-116.  fc [nz, m, n, n]."""
    p = phij[tframes(ntheta, m)]
    return np.einsum('zinm,i->znm', fc, np.conj(p))


def recphi(ftmp, fs, phij, ntheta, m):
    """ftmp_i += fs * phi_j[t_i], kernels_operators.cuh:118-133."""
    p = phij[tframes(ntheta, m)]
    ftmp += fs[:, None, :, :] * p[None, :, None, None]


def mulphi(g, phij, c, m, ntheta):
    """g <- g * ntheta/sqrt(m) * (phi_j[k] or its conjugate), kernels_operators.cuh:135-149."""
    ph = phij if c == 1 else np.conj(phij)
    return g * (ntheta / np.sqrt(m)) * ph[None, :, None]


class Operators:
    """R_apr, R_apr^*, gradient and divergence of one chunk (operators.cu:6-71)."""

    def __init__(self, n, ntheta, m, theta, phi, center, lambda1, dtype=np.float64):
        self.n, self.ntheta, self.m = n, ntheta, m
        self.dtype = np.dtype(dtype)
        self.cdtype = _cdt(dtype)
        self.rad = Usfft(n, theta, center, dtype)
        self.phic = phi_complex(phi).astype(self.cdtype)
        self.lambda1 = lambda1

    # R_apr  (operators.cu:6-32)
    def radonapr(self, f, tau=1.0):
        """h1[z,k,s] = tau * sum_j Re{ntheta/sqrt(m) phi_j[k] fwdR(decphi_j f)}; f [nz, m, n, n] real."""
        nz = f.shape[0]
        g = np.zeros((nz, self.ntheta, self.n), dtype=self.dtype)
        fc = f.astype(self.cdtype)                                   # makecomplex (:21)
        for i in range(self.m):                                      # loop over basis functions (:22)
            fs = decphi(fc, self.phic[i], self.ntheta, self.m)       # (:25)
            gt = self.rad.fwdR(fs)                                   # (:26)
            gt = mulphi(gt, self.phic[i], 1, self.m, self.ntheta)    # (:28)
            g += tau * gt.real                                       # addreal (:30)
        return g

    # R_apr^*  (operators.cu:34-58)
    def radonapradj(self, f, g, tau=1.0):
        """f += tau * Re sum_j phi_j[t_i] adjR(ntheta/sqrt(m) conj(phi_j[k]) g), in place on f."""
        nz = g.shape[0]
        ftmp = np.zeros((nz, self.m, self.n, self.n), dtype=self.cdtype)
        for i in range(self.m):
            gt = g.astype(self.cdtype)                               # makecomplex (:50)
            gt = mulphi(gt, self.phic[i], -1, self.m, self.ntheta)   # (:52)
            fs = self.rad.adjR(gt)                                   # (:53)
            recphi(ftmp, fs, self.phic[i], self.ntheta, self.m)      # (:55)
        f += (tau * ftmp.real).astype(f.dtype)                       # addreal (:57)
        return f

    # gradient  (operators.cu:60-66, kernels_operators.cuh:4-44)
    def gradient(self, fext, flgl, flgr):
        """4D forward differences of the clamped extension.

        fext [nzp+2, m, n, n]: slab 0 = left halo, slabs 1..nzp = the chunk,
        slab nzp+1 = right halo (halo slabs are read only when flgl/flgr = 1).
        Returns h2 [nzp+1, m+1, n+1, n+1, 4] with components (x, y, t, z) in the
        reference's float4 order (.x, .y, .z = time, .w = slice).
        """
        E = extendf(fext, flgl, flgr)
        return gradf(E, self.lambda1)

    def divergent(self, fn, h2, tau):
        """fn += tau * div(h2) (operators.cu:68-71)."""
        return div(fn, h2, tau, self.lambda1)

    def fbp(self, g, r=None):
        """Filtered back-projection of every slice as a static object, the same image in
        each of the m frames: u = pi/(4 sqrt(m)) Re adjR(g, filter=True).  For a basis
        holding the constant function with the others orthogonal to constants at the
        frame times, R_apr of a static object u is sqrt(m) R u (operators.cu:6-32), and
        pi/4 R^*_w inverts R for the applyfilter weights w ~ 4 n |rho| with R, R^*
        scaled by 1/sqrt(n ntheta) (radonusfft.cu:102, :132).  r: optional support mask
        (circ).  Returns [nz, m, n, n]."""
        u = np.pi / (4 * np.sqrt(self.m)) * self.rad.adjR(np.asarray(g).astype(self.cdtype), filter=True).real
        if r is not None:
            u = circ(u, r)
        return np.repeat(u[:, None], self.m, axis=1).astype(self.dtype)


def extendf(fext, flgl, flgr):
    """Clamped/halo extension E [nzp+2, m+2, n+2, n+2], kernels_operators.cuh:4-20."""
    nzp = fext.shape[0] - 2
    m, n = fext.shape[1], fext.shape[2]
    zi = np.clip(np.arange(nzp + 2) - 1, -flgl, nzp - 1 + flgr) + 1
    ti = np.clip(np.arange(m + 2) - 1, 0, m - 1)
    yi = np.clip(np.arange(n + 2) - 1, 0, n - 1)
    xi = np.clip(np.arange(n + 2) - 1, 0, n - 1)
    return fext[np.ix_(zi, ti, yi, xi)]


_CBRT2 = np.cbrt(2.0)


def gradf(E, lambda1):
    """Forward differences, kernels_operators.cuh:22-44.

    .x = dE/dx /2/cbrt(2)/(lambda1/2), .y likewise, .z (time) additionally *lambda1,
    .w (slice) like .x.  Shape [nzp+1, m+1, n+1, n+1, 4].
    """
    s = 1.0 / 2 / _CBRT2 / (lambda1 / 2)
    base = E[:-1, :-1, :-1, :-1]
    h = np.empty(base.shape + (4,), dtype=E.dtype)
    h[..., 0] = (E[:-1, :-1, :-1, 1:] - base) * s
    h[..., 1] = (E[:-1, :-1, 1:, :-1] - base) * s
    h[..., 2] = (E[:-1, 1:, :-1, :-1] - base) * s * lambda1
    h[..., 3] = (E[1:, :-1, :-1, :-1] - base) * s
    return h


def div(fn, h2, tau, lambda1):
    """fn += tau * backward differences of h2, kernels_operators.cuh:46-73."""
    s = 1.0 / 4 / _CBRT2 / (lambda1 / 2)
    c = h2[1:, 1:, 1:, 1:]
    fn += tau * (h2[1:, 1:, 1:, :-1, 0] - c[..., 0]) * s
    fn += tau * (h2[1:, 1:, :-1, 1:, 1] - c[..., 1]) * s
    fn += tau * (h2[1:, :-1, 1:, 1:, 2] - c[..., 2]) * s * lambda1
    fn += tau * (h2[:-1, 1:, 1:, 1:, 3] - c[..., 3]) * s
    return fn


# ---------------------------------------------------------------------------
# ADMM (admm.cu) and its outer loop (rectv.cu)
# ---------------------------------------------------------------------------

def solve_reg(h2, mu, lam, rho):
    """Group soft-threshold of h2 + mu/rho over the 4 components, admm.cu:40-72."""
    psi = h2 + mu / rho
    za = np.sqrt(np.sum(psi * psi, axis=-1, keepdims=True))
    t = lam / rho
    with np.errstate(invalid='ignore', divide='ignore'):
        out = np.where(za <= t, 0.0, psi - t * psi / np.where(za == 0, 1, za))
    return out.astype(h2.dtype)


def solver_admm(ops, fext, g, psi, mu, lambda0, rho, flgl, flgr, titer, stats=None, support=None):
    """One chunk's ADMM step, admm.cu:74-122.

    fext [nzp+2, m, n, n] is the working copy fm with halo slabs; it is updated in
    place.  psi, mu [nzp+1, m+1, n+1, n+1, 4] are updated in place.  Returns
    (||h2 - psi||, ||grad f_old - grad f_new||) of admm.cu:115-117.

    ``stats`` (a dict, optional) receives the two terms of the objective that the
    step evaluates on the way: ``data`` = 1/2 ||R fm - g||^2 of the chunk as it
    enters (the first inner step's residual h1; computed directly when titer = 0)
    and ``tv`` = sum over dual positions of the 4-vector norm of the gradient h2 of
    the updated chunk (the h2 of admm.cu:101, halos from the previous iterate).
    ``support`` (optional r): after each inner step the frames are multiplied by the
    circ mask of radius r (a support constraint; the reference has none).
    """
    inner = fext[1:-1]
    hs = ops.gradient(fext, flgl, flgr)                         # h2stored = grad fm (:79)
    for k in range(titer):
        h2 = ops.gradient(fext, flgl, flgr)                     # (:85)
        h1 = ops.radonapr(inner, 1.0)                           # (:87)
        h2 = rho * h2 - rho * psi + mu                          # lin4 (:90)
        h1 = h1 - g                                             # lin (:92)
        if k == 0 and stats is not None:
            stats["data"] = 0.5 * float(np.sum(h1.astype(np.float64) ** 2))
        ops.divergent(inner, h2, -0.5)                          # (:95)
        ops.radonapradj(inner, h1, -0.5)                        # (:97)
        if support is not None:
            inner *= circ_mask(inner.shape[-1], support)
    if titer == 0 and stats is not None:
        stats["data"] = 0.5 * float(np.sum((ops.radonapr(inner, 1.0) - g).astype(np.float64) ** 2))
    h2 = ops.gradient(fext, flgl, flgr)                         # (:101)
    if stats is not None:
        stats["tv"] = float(np.sum(np.sqrt(np.sum(h2.astype(np.float64) ** 2, axis=-1))))
    psi[...] = solve_reg(h2, mu, lambda0, rho)                  # (:104)
    mu += rho * h2 - rho * psi                                  # (:107)
    hs = hs - h2                                                # (:109)
    h2 = h2 - psi                                               # (:111)
    return float(np.sqrt(np.sum(h2.astype(np.float64) ** 2))), float(np.sqrt(np.sum(hs.astype(np.float64) ** 2)))


def run(data, theta, phi, n, ntheta, m, nz, nzp, center, lambda0, lambda1, niter, titer,
        dtype=np.float64, log=None, history=None, order=None, state=None, threads=1, tol=None,
        support=None):
    """rectv::run, rectv.cu:102-252 (lambda0 arrives already multiplied by lambda1).

    Chunks of nzp slices are processed against the previous outer iterate f (halo
    slabs included) and write fn; then rho is rebalanced (r > 10 s -> rho*2,
    s > 10 r -> rho/2, rectv.cu:224-232), ||fn - f||^2 is reported and f/fn swap.

    ``history`` (a list) receives one dict per outer iteration: r, s, rho used,
    rho for the next iteration, rdiff = ||fn - f||^2, ||fn||^2, and the objective
    terms ``data`` = 1/2 ||R_apr f - g||^2 of the iterate entering the iteration
    and ``tv`` = sum of the dual-position gradient norms of the iterate leaving it
    (per chunk, halos from the entering iterate, ghost positions included; see
    solver_admm).  The fixed point of the iteration minimises (up to the ghost
    positions) J(f) = 1/2 ||R_apr f - g||^2 + lambda0/2 * sum ||grad f||: the inner
    step descends 1/2||R f - g||^2 + 1/2 (<mu, grad f - psi> + rho/2 ||grad f - psi||^2)
    because div is half the transpose of grad, i.e. ADMM with penalty rho/2 on
    lambda0/2 ||psi||_{2,1}.
    ``order`` (a permutation of the chunk indices, default ascending) sets the
    order in which chunks are processed; the reference reads f and writes fn, so
    the result does not depend on it (rectv.cu:170-171).
    ``state`` (a dict, optional): entries f, psi, mu, rho present on entry replace
    the reference's zero start (rectv.cu:119-125); on return it holds the final
    f, psi, mu and rho.
    ``threads`` > 1 processes that many chunks at a time on host threads (chunks
    read f and write disjoint parts of fn, psi and mu, so the result is the same).
    ``tol``: stop after the first outer iteration whose relative change
    sqrt(||fn - f||^2 / ||fn||^2) is below tol (the early stop of the serving loop;
    the reference always runs niter iterations).
    ``support``: the circ-mask radius parameter of solver_admm (None: no mask).
    Returns f [nz, m, n, n].
    """
    if nz % nzp != 0:
        raise ValueError("nz must be a multiple of nzp")
    dt = np.dtype(dtype)
    nch = nz // nzp
    ops = Operators(n, ntheta, m, theta, phi, center, lambda1, dt)
    g = np.asarray(data, dtype=np.float32).astype(dt).reshape(nz, ntheta, n)
    st = {} if state is None else state
    f = np.array(st["f"], dt) if "f" in st else np.zeros((nz, m, n, n), dt)
    fn = np.zeros_like(f)
    shp = (nch, nzp + 1, m + 1, n + 1, n + 1, 4)
    psi = np.array(st["psi"], dt).reshape(shp) if "psi" in st else np.zeros(shp, dt)
    mu = np.array(st["mu"], dt).reshape(shp) if "mu" in st else np.zeros(shp, dt)
    rho = float(st.get("rho", 0.5))
    chunks = list(range(nch)) if order is None else [int(c) for c in order]
    if sorted(chunks) != list(range(nch)):
        raise ValueError("order must be a permutation of the chunk indices")
    pool = None
    if threads > 1:
        import concurrent.futures as cf
        pool = cf.ThreadPoolExecutor(max_workers=threads)
    for it in range(niter):
        nd = [None] * nch
        sts = [dict() for _ in range(nch)]

        def chunk(iz):
            flgl, flgr = int(iz != 0), int(iz != nch - 1)
            fext = np.zeros((nzp + 2, m, n, n), dt)
            lo = iz * nzp
            fext[1:-1] = f[lo:lo + nzp]
            if flgl:
                fext[0] = f[lo - 1]
            if flgr:
                fext[-1] = f[lo + nzp]
            nd[iz] = solver_admm(ops, fext, g[lo:lo + nzp], psi[iz], mu[iz],
                                 lambda0, rho, flgl, flgr, titer, stats=sts[iz], support=support)
            fn[lo:lo + nzp] = fext[1:-1]

        if pool is None:
            for iz in chunks:
                chunk(iz)
        else:
            list(pool.map(chunk, chunks))
        data_t = sum(x["data"] for x in sts)
        tv_t = sum(x["tv"] for x in sts)
        r = sum(x * x for x, _ in nd)
        s = sum(rho * rho * y * y for _, y in nd)
        rho_used = rho
        if r > 10 * s:
            rho *= 2
        elif s > 10 * r:
            rho *= 0.5
        norm = float(np.sum((fn.astype(np.float64) - f.astype(np.float64)) ** 2))
        if log is not None:
            print("iter (%d/%d), rdiff ||f(k+1)-f(k)||:%f " % (it, niter, norm), file=log)
        if history is not None:
            history.append(dict(iter=it, r=r, s=s, rho=rho_used, rho_next=rho, rdiff=norm,
                                fnorm=float(np.sum(fn.astype(np.float64) ** 2)), data=data_t, tv=tv_t))
        f, fn = fn, f
        fn2 = float(np.sum(f.astype(np.float64) ** 2))
        if tol is not None and fn2 > 0 and np.sqrt(norm / fn2) < tol:
            break
    if pool is not None:
        pool.shutdown()
    if state is not None:
        state.update(f=f, psi=psi, mu=mu, rho=rho)
    return f


def rho_margin(r, s):
    """Distance of an outer iteration's (r, s) from the two thresholds of the rho
    rule (rectv.cu:231-232) as min(|log(r / 10 s)|, |log(s / 10 r)|): near zero, FP32
    and FP64 arithmetic may take different branches (inf when r or s is zero)."""
    if r <= 0 or s <= 0:
        return np.inf
    return float(min(abs(np.log(r / (10 * s))), abs(np.log(s / (10 * r)))))


def check_adjoints(data, theta, phi, n, ntheta, m, nz, nzp, lambda1, center, dtype=np.float64):
    """rectv::check_adjoints, rectv.cu:255-289 (valid for nz == nzp).

    res[0] = ||R* g||^2, res[1] = <g, R R* g>, res[2] = ||R R* g||^2,
    res[3] = <f, div grad f> (f = R* g), res[4] = sum of squares of the first
    (n+1)^2 (m+1)(nz+1) floats of the float4 gradient array, res[5] = ||div grad f||^2.
    """
    if nz != nzp:
        raise ValueError("check_adjoints requires nz == nzp (one chunk), as the reference states")
    dt = np.dtype(dtype)
    ops = Operators(n, ntheta, m, theta, phi, center, lambda1, dt)
    g = np.asarray(data, dtype=np.float32).astype(dt).reshape(nz, ntheta, n)
    f = np.zeros((nz, m, n, n), dt)
    ops.radonapradj(f, g, 1.0)
    h1 = ops.radonapr(f, 1.0)
    res = np.zeros(6)
    res[0] = np.sum(f.astype(np.float64) ** 2)
    res[1] = np.sum(g.astype(np.float64) * h1)
    res[2] = np.sum(h1.astype(np.float64) ** 2)
    fext = np.zeros((nz + 2, m, n, n), dt)
    fext[1:-1] = f
    h2 = ops.gradient(fext, 0, 0)
    fn = np.zeros_like(f)
    ops.divergent(fn, h2, 1.0)
    q = (n + 1) * (n + 1) * (m + 1) * (nz + 1)
    res[3] = np.sum(f.astype(np.float64) * fn)
    res[4] = np.sum(h2.reshape(-1)[:q].astype(np.float64) ** 2)
    res[5] = np.sum(fn.astype(np.float64) ** 2)
    return res


# ---------------------------------------------------------------------------
# Inputs used by the reference's own test and examples
# ---------------------------------------------------------------------------

def takephi(ntheta, m):
    """Fourier temporal basis, tests/test_adjoint.py:23-27 and foam_example.py:17-23.

    phi_j[k] = exp(2 pi i a_k b_j/ntheta)/sqrt(ntheta) with a = arange(-ntheta//2,
    ntheta//2), b = arange(-m//2, m//2) (Python floor division, so odd sizes start
    at -(size+1)/2), stored as float32 [m, 2*ntheta] with re/im interleaved.
    """
    x, y = np.meshgrid(np.arange(-ntheta // 2, ntheta // 2), np.arange(-m // 2, m // 2))
    phi = np.zeros([m, 2 * ntheta], dtype='float32')
    phi[:, ::2] = np.cos(2 * np.pi * x * y / ntheta) / np.sqrt(ntheta)
    phi[:, 1::2] = np.sin(2 * np.pi * x * y / ntheta) / np.sqrt(ntheta)
    return phi


if __name__ == "__main__":  # tiny self-run
    n, nt, m, nz = 32, 16, 4, 2
    th = (np.arange(nt) * np.pi / 8).astype(np.float32)
    ph = takephi(nt, m)
    d = np.random.default_rng(0).random((nz, nt, n)).astype(np.float32)
    print(check_adjoints(d, th, ph, n, nt, m, nz, nz, 4.0, n / 2))
    print(run(d, th, ph, n, nt, m, nz, 1, n / 2, 1e-3 * 4, 4.0, 2, 2, log=sys.stderr).shape)


# ---------------------------------------------------------------------------
# Pre-processing before the solver (examples/tao/tao.py, examples/tao/rec_line.txt)
# ---------------------------------------------------------------------------

def darkflat_log(data, dark, flat):
    """-log((data - mean(dark)) / (1e-3 + mean(flat) - mean(dark))) with the frame
    means over axis 0, then swapaxes(0, 1) to sinograms (examples/tao/tao.py:34-43).
    data [ntheta, nz, n], dark [ndark, nz, n], flat [nflat, nz, n] -> [nz, ntheta, n]."""
    d = np.asarray(data, np.float64)
    dm = np.mean(np.asarray(dark, np.float64), axis=0)
    fm = np.mean(np.asarray(flat, np.float64), axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        out = -np.log((d - dm) / (1e-3 + fm - dm))
    return np.ascontiguousarray(out.swapaxes(0, 1))


def _py_slice_bounds(lo, hi, n):
    """Bounds of a[lo:hi] for a length-n axis under Python's slice rules."""
    return slice(lo, hi).indices(n)[:2]


def remove_stripe_ti(g, beta, mask_size):
    """Stripe removal of examples/tao/tao.py:58-71 (after V. Titarenko) on sinograms
    g [nz, ntheta, n] (the example applies it to data.swapaxes(0, 1), i.e. with the
    mean taken over the projections): gamma_k = beta ((1-beta)/(1+beta))^|k| over the
    signed detector frequencies k (gamma_0 - 1), v = mean over projections minus its
    first column, v <- v (*) gamma (circular convolution via rfft/irfft), and v is
    added to every projection on the columns [n//2 - ms//2, n//2 + ms//2) with
    ms = mask_size * n (integer mask_size, Python slice rules).  Even n."""
    g = np.array(g, np.float64)
    n = g.shape[-1]
    gamma = beta * ((1 - beta) / (1 + beta)) ** np.abs(np.fft.fftfreq(n) * n)
    gamma[0] -= 1
    v = np.mean(g, axis=1)
    v = v - v[:, 0:1]
    v = np.fft.irfft(np.fft.rfft(v) * np.fft.rfft(gamma), n)
    ms = int(mask_size) * n
    lo, hi = _py_slice_bounds(n // 2 - ms // 2, n // 2 + ms // 2, n)
    mask = np.zeros(n)
    mask[lo:hi] = 1
    return g + (v * mask)[:, None, :]


# Physical constants of tomopy's phase retrieval (tomopy/prep/phase.py, tomopy 1.x)
_PLANCK_KEV_S = 6.58211928e-19  # reduced Planck constant [keV s]
_SPEED_OF_LIGHT_CM = 299792458e+2  # [cm/s]
_PI_TOMOPY = 3.14159265359


def wavelength_cm(energy_kev):
    """2 pi hbar c / E [cm] (tomopy's _wavelength)."""
    return 2 * _PI_TOMOPY * _PLANCK_KEV_S * _SPEED_OF_LIGHT_CM / energy_kev


def paganin_pad_width(dim, pixel_size, dist, energy):
    """Pad width of tomopy's _calc_pad_width: half the distance to the next power of two
    above dim + ceil(pi lambda dist / pixel^2)."""
    pad_pix = np.ceil(_PI_TOMOPY * wavelength_cm(energy) * dist / pixel_size ** 2)
    return int((pow(2, np.ceil(np.log2(dim + pad_pix))) - dim) * 0.5)


def _reciprocal_coord(pixel_size, num_grid):
    n = num_grid - 1
    rc = np.arange(-n, num_grid, 2, dtype=np.float64)
    return rc * (2 * _PI_TOMOPY / (n * pixel_size))


def paganin_filter(ny, nz, pixel_size, dist, energy, alpha):
    """Normalised single-material phase filter (Paganin et al., J. Microsc. 206 (2002)
    33-40) as tomopy's retrieve_phase forms it on a padded ny x nz projection:
    H = 1/(lambda dist w2/(4 pi) + alpha) on the grid w2 = kx^2 + ky^2 of
    _reciprocal_grid, fftshift-ed and divided by its maximum."""
    w2 = np.add.outer(_reciprocal_coord(pixel_size, ny) ** 2, _reciprocal_coord(pixel_size, nz) ** 2)
    h = np.fft.fftshift(1.0 / (wavelength_cm(energy) * dist * w2 / (4 * _PI_TOMOPY) + alpha))
    return h / h.max()


def retrieve_phase(proj, pixel_size=1e-4, dist=50.0, energy=20.0, alpha=1e-3):
    """Paganin phase retrieval of every projection proj[k] (proj [ntheta, ny, nz],
    intensities after dark/flat normalisation), restating tomopy's retrieve_phase with
    pad=True: the projection is edge-extended by the pad widths of paganin_pad_width on
    each side, multiplied in 2D Fourier space by paganin_filter and cropped back.
    Units: pixel_size and dist in cm, energy in keV (tomocupy's command line of
    examples/tao/rec_line.txt:1 gives 3.13 um, 250 mm, 25 keV, alpha 5e-4)."""
    proj = np.asarray(proj, np.float64)
    nt, ny, nz = proj.shape
    py = paganin_pad_width(ny, pixel_size, dist, energy)
    pz = paganin_pad_width(nz, pixel_size, dist, energy)
    H = paganin_filter(ny + 2 * py, nz + 2 * pz, pixel_size, dist, energy, alpha)
    out = np.empty_like(proj)
    for k in range(nt):
        p = np.pad(proj[k], ((py, py), (pz, pz)), mode="edge")
        out[k] = np.real(sfft.ifft2(sfft.fft2(p) * H))[py:py + ny, pz:pz + nz]
    return out
