"""Pins for the oracle's FBP filter, support mask and filtered back-projection
(kernels_radonusfft.cuh:146-189, the unused `filter` argument of radonusfft::adjR,
radonusfft.cu:106): a hand-counted mask, the filter's shape, the filter's place
in adjR checked against an explicit circular convolution on the detector, the
reconstruction of an analytic disk, and the warm start's effect on ADMM."""
import numpy as np
import scipy.fft as sfft

from oracle import rectv_oracle as O


def test_circ_mask_known_answer():
    # n = 8, r = 0: x = (tx - 4)/8, kept where 4x^2 + 4y^2 < 1; counted by hand per row
    m = O.circ_mask(8, 0.0)
    assert list(m.sum(axis=1).astype(int)) == [0, 5, 7, 7, 7, 7, 7, 5]
    assert m[4, 1] == 1 and m[4, 0] == 0 and m[1, 2] == 1 and m[1, 1] == 0
    np.testing.assert_array_equal(m[1:, 1:], m[1:, 1:][::-1, ::-1])  # symmetric about (4, 4)
    # a larger r shrinks the support (r = 0.5: 4x^2 + 4y^2 < 0.5, i.e. (tx-4)^2+(ty-4)^2 < 8)
    m5 = O.circ_mask(8, 0.5)
    assert m5.sum() == 21 and np.all(m5 <= m)


def test_filter_shape():
    n = 64
    w = O.applyfilter_weights(n)
    rho = (np.arange(n) - n // 2) / n
    assert w[n // 2] == 0
    np.testing.assert_allclose(w[n // 2 + 1:], w[n // 2 - 1:0:-1])            # even in rho
    assert np.all(np.diff(w[n // 2:]) > 0)                                      # rising with |rho|
    np.testing.assert_allclose(w[1:], 4 * n * np.abs(rho[1:]), rtol=0.5 ** 2 / 6 + 1e-12)  # ~ ramp 4n|rho|


def test_filter_is_a_detector_convolution_before_the_adjoint():
    """adjR(g, filter=True) == adjR(h * g) with h the circular convolution whose DFT is
    the filter on the centred spectrum (frequency index k <-> centred index (k + n/2) mod n)."""
    n, nt = 32, 12
    th = np.sort(np.random.default_rng(3).uniform(0, 2 * np.pi, nt)).astype(np.float32)
    U = O.Usfft(n, th, n / 2 + 0.75)
    g = np.random.default_rng(4).standard_normal((2, nt, n)) + 0j
    W = O.applyfilter_weights(n)
    Wk = W[(np.arange(n) + n // 2) % n]
    gf = sfft.ifft(sfft.fft(g, axis=-1) * Wk, axis=-1)
    a = U.adjR(g, filter=True)
    b = U.adjR(gf)
    assert np.linalg.norm(a - b) < 1e-12 * np.linalg.norm(b)


def _disk_case(n=64, nrot=4, ntr=64, m=4):
    th = (np.arange(nrot * ntr) * np.pi / ntr).astype(np.float32)
    nt = len(th)
    phi = O.takephi(nt, m)
    yy, xx = np.mgrid[0:n, 0:n]
    d2 = (xx - n / 2 - 5) ** 2 + (yy - n / 2 + 3) ** 2
    u = (d2 < 15 ** 2).astype(float)
    ops = O.Operators(n, nt, m, th, phi, n / 2, 4.0)
    g = ops.radonapr(np.broadcast_to(u, (1, m, n, n)).copy())
    return n, nt, m, th, phi, ops, g, d2


def test_fbp_reconstructs_a_static_disk():
    """pi/(4 sqrt(m)) Re R^*_w of the data of a static disk of density 1 is flat inside
    the disk at 0.91: the filter's zero weight at zero frequency (no padding) drops the
    image mean, which offsets the whole image by about -0.1 (outside the disk too)."""
    n, nt, m, th, phi, ops, g, d2 = _disk_case()
    f = ops.fbp(g)
    assert f.shape == (1, m, n, n) and np.all(f[0, 0] == f[0, m - 1])
    inside, outside = d2 < 12 ** 2, (d2 > 18 ** 2) & (O.circ_mask(n, 0.05) > 0)
    mi, mo = f[0, 0][inside].mean(), f[0, 0][outside].mean()
    assert 0.85 < mi < 0.95 and f[0, 0][inside].std() < 0.05
    assert 0.9 < mi - mo < 1.1   # the disk's contrast is 1
    fm = ops.fbp(g, r=0.05)
    assert not fm[0, 0][O.circ_mask(n, 0.05) == 0].any()


def test_fbp_warm_start_lowers_the_objective():
    n, nt, m, th, phi, ops, g, d2 = _disk_case(n=32, nrot=2, ntr=32, m=2)
    g = g.astype(np.float32)
    lam0, lam1 = 1e-3 * 4, 4.0
    for niter in (1, 3):
        h0, h1 = [], []
        O.run(g, th, phi, n, nt, m, 1, 1, n / 2, lam0, lam1, niter + 1, 2, history=h0)
        O.run(g, th, phi, n, nt, m, 1, 1, n / 2, lam0, lam1, niter + 1, 2, history=h1, state={"f": ops.fbp(g)})
        J0 = h0[niter]["data"] + lam0 / 2 * h0[niter - 1]["tv"]
        J1 = h1[niter]["data"] + lam0 / 2 * h1[niter - 1]["tv"]
        assert J1 < 0.5 * J0, (niter, J0, J1)


def test_support_mask_in_the_iterations():
    n, nt, m, th, phi, ops, g, d2 = _disk_case(n=32, nrot=2, ntr=32, m=2)
    f = O.run(g.astype(np.float32), th, phi, n, nt, m, 1, 1, n / 2, 4e-3, 4.0, 2, 2, support=0.1)
    assert not f[..., O.circ_mask(n, 0.1) == 0].any() and f.any()
