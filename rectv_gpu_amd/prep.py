"""Pre-processing of projection data before the solver, on the GPU (C ABI of
include/rectv_amd.h).  The reference leaves these steps to its examples and to
tomocupy (examples/tao/tao.py:34-43, 58-71; examples/tao/rec_line.txt:1):

  * ``darkflat(data, dark, flat)``: -log((data - mean dark) / (1e-3 + mean flat -
    mean dark)), returned as sinograms [nz, ntheta, n] (tao.py:34-43);
  * ``remove_stripe_ti(g, beta, mask_size)``: the stripe removal of tao.py:58-71 on
    sinograms [nz, ntheta, n];
  * ``retrieve_phase(proj, pixel_size, dist, energy, alpha)``: single-material
    (Paganin) phase retrieval of projections [ntheta, ny, nz] as tomopy's
    retrieve_phase applies it (pixel_size, dist in cm; energy in keV).
"""
import numpy as np

from . import _lib
from .solver import getp


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def darkflat(data, dark, flat, log=True):
    """data [ntheta, nz, n], dark [ndark, nz, n], flat [nflat, nz, n] -> [nz, ntheta, n] f32."""
    data, dark, flat = _f32(data), _f32(dark), _f32(flat)
    if data.ndim != 3 or dark.ndim != 3 or flat.ndim != 3 or dark.shape[1:] != data.shape[1:] \
            or flat.shape[1:] != data.shape[1:]:
        raise ValueError("data, dark, flat must be [frames, nz, n] with equal (nz, n)")
    nt, nz, n = data.shape
    out = np.empty((nz, nt, n), np.float32)
    _lib.check(_lib.lib().rtv_prep_darkflat(getp(data), getp(dark), getp(flat), nt, nz, n, dark.shape[0],
                                            flat.shape[0], 1 if log else 0, getp(out)))
    return out


def remove_stripe_ti(g, beta=0.022, mask_size=1):
    """Stripe removal of sinograms g [nz, ntheta, n] (returns a new array)."""
    g = _f32(g).copy()
    if g.ndim != 3:
        raise ValueError("g must be [nz, ntheta, n]")
    nz, nt, n = g.shape
    _lib.check(_lib.lib().rtv_prep_stripe_ti(getp(g), nz, nt, n, float(beta), int(mask_size)))
    return g


def retrieve_phase(proj, pixel_size=1e-4, dist=50.0, energy=20.0, alpha=1e-3):
    """Paganin phase retrieval of projections proj [ntheta, ny, nz] (returns a new array)."""
    p = _f32(proj).copy()
    if p.ndim != 3:
        raise ValueError("proj must be [ntheta, ny, nz]")
    nt, ny, nz = p.shape
    _lib.check(_lib.lib().rtv_prep_paganin(getp(p), nt, ny, nz, float(pixel_size), float(dist), float(energy),
                                           float(alpha)))
    return p
