"""Pins for the oracle's pre-processing (examples/tao/tao.py:34-43, 58-71; phase
retrieval of examples/tao/rec_line.txt:1): known answers built from the inverse of
the normalisation, the stripe removal's exact response to a single stripe (it
replaces the projection-mean profile by its convolution with the normalised
geometric kernel gamma), and the phase filter's unit gain at zero frequency and
its limit for large alpha.  tomopy / tomocupy are not available here, so the phase
retrieval is pinned by these properties only."""
import numpy as np

from oracle import rectv_oracle as O


def test_darkflat_inverts_a_known_attenuation():
    rng = np.random.default_rng(1)
    nt, nz, n = 5, 3, 8
    p = rng.uniform(0.1, 3.0, (nt, nz, n))
    dark = rng.uniform(90, 110, (2, nz, n))
    flat = rng.uniform(900, 1100, (3, nz, n))
    dm, fm = dark.mean(axis=0), flat.mean(axis=0)
    data = np.exp(-p) * (1e-3 + fm - dm) + dm
    out = O.darkflat_log(data, dark, flat)
    assert out.shape == (nz, nt, n)
    np.testing.assert_allclose(out, p.swapaxes(0, 1), rtol=1e-12)


def test_stripe_removal_response_to_one_stripe():
    """A stripe of height A in column c (constant over projections) on a background
    constant along the detector becomes A * beta * q^|s - c| with q = (1-beta)/(1+beta)
    (circular distance): the removal swaps the mean profile for its smoothing."""
    nz, nt, n, c, A, beta = 2, 7, 32, 9, 3.0, 0.022
    rng = np.random.default_rng(2)
    bg = rng.uniform(0, 1, (nz, nt, 1)) * np.ones((1, 1, n))
    g = bg.copy()
    g[:, :, c] += A
    out = O.remove_stripe_ti(g, beta, 1)
    q = (1 - beta) / (1 + beta)
    d = np.abs(np.arange(n) - c)
    d = np.minimum(d, n - d)
    np.testing.assert_allclose(out - bg, np.broadcast_to(A * beta * q ** d, out.shape), atol=1e-12)
    np.testing.assert_array_equal(O.remove_stripe_ti(g, beta, 0), g)  # mask_size 0: empty mask


def test_phase_filter_gain_and_limits():
    ny, nz, px, dist, en = 24, 40, 3.13e-4, 25.0, 25.0
    H = O.paganin_filter(ny, nz, px, dist, en, 5e-4)
    assert H.max() == 1.0 and H[0, 0] == 1.0                 # unit gain at the lowest frequency
    assert np.all(H > 0) and np.all(H <= 1)
    rng = np.random.default_rng(3)
    const = np.full((2, ny, nz), 0.7)
    np.testing.assert_allclose(O.retrieve_phase(const, px, dist, en, 5e-4), const, rtol=1e-12)
    proj = rng.uniform(0.5, 1.0, (2, ny, nz))
    near_id = O.retrieve_phase(proj, px, dist, en, 1e12)      # alpha -> infinity: filter -> 1
    np.testing.assert_allclose(near_id, proj, rtol=1e-9)
    smooth = O.retrieve_phase(proj, px, dist, en, 5e-4)       # a low-pass: less variation
    assert np.std(smooth) < np.std(proj)
    np.testing.assert_allclose(smooth.mean(axis=(1, 2)), proj.mean(axis=(1, 2)), rtol=0.05)
