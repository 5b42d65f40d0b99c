"""Input helpers used by the reference's test and examples."""
import numpy as np


def takephi(ntheta, m):
    """Fourier temporal basis as float32 [m, 2*ntheta] (re/im interleaved):
    phi_j[k] = exp(2 pi i a_k b_j / ntheta)/sqrt(ntheta) with a = -ntheta//2, ...,
    ntheta//2 - 1 and b = -m//2, ..., m//2 - 1 in Python floor division, i.e.
    a_k = k + floor(-ntheta/2) (tests/test_adjoint.py:23-27,
    examples/foam/foam_example.py:17-23; for odd sizes the ranges start at
    -(size+1)/2)."""
    k = np.arange(-ntheta // 2, ntheta // 2)
    j = np.arange(-m // 2, m // 2)
    x, y = np.meshgrid(k, j)
    phi = np.zeros([m, 2 * ntheta], dtype='float32')
    phi[:, ::2] = np.cos(2 * np.pi * x * y / ntheta) / np.sqrt(ntheta)
    phi[:, 1::2] = np.sin(2 * np.pi * x * y / ntheta) / np.sqrt(ntheta)
    return phi


def angles(ntheta_per_rot, nrot):
    """theta_k = k*pi/ntheta_per_rot for k < nrot*ntheta_per_rot (float32), i.e. nrot
    half-turns of ntheta_per_rot projections each."""
    k = np.arange(nrot * ntheta_per_rot, dtype=np.float64)
    return (k * np.pi / ntheta_per_rot).astype(np.float32)


def realbasis(ntheta, m):
    """Real Fourier basis over the acquisition {1, cos, sin, cos2, sin2, ...}, normalised
    to unit l2 norm over the ntheta projections, as float32 [m, 2*ntheta] with zero
    imaginary parts."""
    k = np.arange(ntheta, dtype=np.float64)
    phi = np.zeros([m, 2 * ntheta], dtype='float32')
    for j in range(m):
        if j == 0:
            b = np.ones(ntheta) / np.sqrt(ntheta)
        else:
            h = (j + 1) // 2
            f = np.cos if j % 2 == 1 else np.sin
            b = f(2 * np.pi * h * k / ntheta) * np.sqrt(2.0 / ntheta)
        phi[j, ::2] = b
    return phi
