"""Pins for the FP64 oracle: each check is independent of the oracle's own formulas
(explicit DFT matrices, brute-force nonuniform DFT, analytic line integrals,
summation by parts, numerical minimisation, invariants) so that a dropped term, a
wrong sign or index, or a transposed operand in the oracle fails one of them.
"""
import numpy as np
import pytest
import scipy.optimize

from oracle import rectv_oracle as O


def _angles(nt, seed=1, span=8 * np.pi):
    rng = np.random.default_rng(seed)
    return np.sort(rng.uniform(0, span, nt)).astype(np.float32)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ---------------------------------------------------------------- USFFT geometry
@pytest.mark.parametrize("n,nt", [(32, 20), (48, 16)])
def test_fwdR_matches_bruteforce_nudft(n, nt):
    """fwdR == brute-force NUDFT at the polar points followed by the 1D stage written
    with an explicit inverse-DFT matrix (definition, e^{+2 pi i s s'/n}).  The only
    approximation is the USFFT gridding itself (eps = 1e-3)."""
    th = _angles(nt)
    center = n / 2 + 1.5
    U = O.Usfft(n, th, center)
    rng = np.random.default_rng(3)
    u = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    x0, y0 = O.takexy(th, n)
    xc = np.arange(n) - n // 2
    # F[b,k,s] = sum_{y,x} u[b,y,x] exp(-2 pi i ((x-n/2) x0 + (y-n/2) y0))
    ex = np.exp(-2j * np.pi * xc[None, None, :] * x0[:, :, None])     # [k,s,x]
    ey = np.exp(-2j * np.pi * xc[None, None, :] * y0[:, :, None])     # [k,s,y]
    F = np.einsum('byx,ksx,ksy->bks', u, ex, ey)
    s = np.arange(n)
    g = F / n
    g[:, :, 0] = 0
    c = np.float32(center - n / 2.0)
    g = g * np.exp(-2j * np.pi * float(c) * (s - n / 2.0) / n)[None, None, :]
    sig = 1 - 2 * ((s + 1) % 2)
    Dinv = np.exp(2j * np.pi * np.outer(s, s) / n)
    g = ((g * sig) @ Dinv) * sig / np.sqrt(n * nt)
    out = U.fwdR(u)
    assert _rel(out, g) < 2e-3


def test_fwdR_gaussian_blob_line_integrals():
    """Fourier-slice check: fwdR of a Gaussian blob equals its analytic line integrals
    sqrt(2 pi) sigma exp(-(s_c - p_k)^2 / 2 sigma^2) / sqrt(n ntheta), with
    p_k = (xb - n/2) sin(theta) + (yb - n/2) cos(theta) and the rotation centre at
    detector position ``center``."""
    n, nt = 64, 24
    th = _angles(nt, seed=5)
    sig, xb, yb = 2.5, 40.0, 25.0
    yy, xx = np.mgrid[0:n, 0:n]
    u = np.exp(-((xx - xb) ** 2 + (yy - yb) ** 2) / (2 * sig ** 2)).astype(complex)
    thd = th.astype(np.float64)
    p = (xb - n // 2) * np.sin(thd) + (yb - n // 2) * np.cos(thd)
    for shift in (0.0, 3.0):
        U = O.Usfft(n, th, n / 2 + shift)
        g = U.fwdR(u[None])[0] * np.sqrt(n * nt)
        sc = np.arange(n) - n // 2 - shift
        P = np.sqrt(2 * np.pi) * sig * np.exp(-(sc[None, :] - p[:, None]) ** 2 / (2 * sig ** 2))
        assert np.abs(g - P).max() / P.max() < 5e-5


def test_usfft_adjointness():
    n, nt = 40, 18
    th = _angles(nt, seed=7)
    U = O.Usfft(n, th, n / 2 - 2.25)
    rng = np.random.default_rng(4)
    u = rng.standard_normal((3, n, n)) + 1j * rng.standard_normal((3, n, n))
    v = rng.standard_normal((3, nt, n)) + 1j * rng.standard_normal((3, nt, n))
    lhs = np.vdot(v, U.fwdR(u))
    rhs = np.vdot(U.adjR(v), u)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_usfft_constants():
    for n in (32, 128, 504, 1024):
        mu, K = O.usfft_params(n)
        assert K == 4
        assert abs(mu * 2 * n * n - (-np.log(1e-3))) < 1e-5


# ---------------------------------------------------------------- basis operators
def test_decphi_is_centred_dft_across_frames():
    nt, m, n = 64, 8, 6
    phic = O.phi_complex(O.takephi(nt, m))
    rng = np.random.default_rng(2)
    f = rng.standard_normal((2, m, n, n))
    F = np.fft.fft(f, axis=1)                          # sum_i f_i e^{-2 pi i i q/m}
    for j in range(m):
        q = (j - m // 2) % m
        expect = (-1.0) ** (j - m // 2) * F[:, q] / np.sqrt(nt)
        got = O.decphi(f.astype(complex), phic[j], nt, m)
        assert np.allclose(got, expect, atol=1e-6 * np.abs(expect).max())


def test_single_constant_basis_reduces_to_plain_radon():
    n, nt = 32, 20
    th = _angles(nt, seed=9)
    phi = np.zeros((1, 2 * nt), np.float32)
    phi[0, 0::2] = 1 / np.sqrt(nt)
    ops = O.Operators(n, nt, 1, th, phi, n / 2, 1.0)
    rng = np.random.default_rng(8)
    f = rng.standard_normal((2, 1, n, n))
    h = ops.radonapr(f, 1.0)
    ref = ops.rad.fwdR(f[:, 0].astype(complex)).real * nt / np.sqrt(nt) * np.float64(np.float32(1 / np.sqrt(nt)))
    assert _rel(h, ref) < 1e-7


def test_radonapr_adjointness_fourier_basis():
    n, nt, m = 32, 32, 4
    th = (np.arange(nt) * np.pi / 8).astype(np.float32)
    ops = O.Operators(n, nt, m, th, O.takephi(nt, m), n / 2 + 0.5, 2.0)
    rng = np.random.default_rng(11)
    f = rng.standard_normal((2, m, n, n))
    g = rng.standard_normal((2, nt, n))
    Rf = ops.radonapr(f, 1.0)
    Rsg = ops.radonapradj(np.zeros_like(f), g, 1.0)
    assert abs(np.vdot(g, Rf) - np.vdot(Rsg, f)) / abs(np.vdot(g, Rf)) < 1e-12


# ---------------------------------------------------------------- gradient / divergence
def test_gradient_matches_np_diff_with_clamped_borders():
    nzp, m, n, lam = 3, 4, 5, 3.0
    rng = np.random.default_rng(12)
    fext = rng.standard_normal((nzp + 2, m, n, n))
    for flgl, flgr in ((0, 0), (1, 0), (0, 1), (1, 1)):
        h = O.gradf(O.extendf(fext, flgl, flgr), lam)
        core = fext[1:-1]
        lo = fext[:1] if flgl else core[:1]
        hi = fext[-1:] if flgr else core[-1:]
        P = np.concatenate([lo, core, hi], axis=0)
        P = np.pad(P, ((0, 0), (1, 1), (1, 1), (1, 1)), mode='edge')
        c = 1 / np.cbrt(2.0) / lam
        assert np.allclose(h[..., 0], np.diff(P, axis=3)[:-1, :-1, :-1, :] * c)
        assert np.allclose(h[..., 1], np.diff(P, axis=2)[:-1, :-1, :, :-1] * c)
        assert np.allclose(h[..., 2], np.diff(P, axis=1)[:-1, :, :-1, :-1] / np.cbrt(2.0))
        assert np.allclose(h[..., 3], np.diff(P, axis=0)[:, :-1, :-1, :-1] * c)


def test_divergence_is_half_adjoint_of_gradient():
    """Summation by parts: <f, div h> = +1/2 <grad f, h> for h vanishing on the border
    dual planes (h = grad g has this property), single chunk, Neumann borders.  The
    reference's div takes h[v-1] - h[v] (kernels_operators.cuh:63-72), i.e. the
    transpose of its forward difference, at half the gradient's scale."""
    nzp, m, n, lam = 3, 4, 6, 2.5
    rng = np.random.default_rng(13)
    f = rng.standard_normal((nzp, m, n, n))
    g = rng.standard_normal((nzp, m, n, n))
    ext = lambda a: np.concatenate([a[:1], a, a[-1:]], axis=0)
    hf = O.gradf(O.extendf(ext(f), 0, 0), lam)
    hg = O.gradf(O.extendf(ext(g), 0, 0), lam)
    dg = O.div(np.zeros_like(f), hg, 1.0, lam)
    # the entries div reads: component c on planes >= 1 of the other three axes
    mask = np.zeros_like(hf)
    mask[1:, 1:, 1:, :, 0] = 1
    mask[1:, 1:, :, 1:, 1] = 1
    mask[1:, :, 1:, 1:, 2] = 1
    mask[:, 1:, 1:, 1:, 3] = 1
    lhs = np.sum(f * dg)
    rhs = 0.5 * np.sum(hf * hg * mask)
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)


# ---------------------------------------------------------------- prox / ADMM
def test_solve_reg_is_group_soft_threshold_minimiser():
    rng = np.random.default_rng(14)
    lam, rho = 0.8, 0.5
    for _ in range(6):
        h = rng.standard_normal(4) * rng.uniform(0.2, 2.0)
        mu = rng.standard_normal(4) * 0.3
        v = h + mu / rho
        t = lam / rho
        obj = lambda p: 0.5 * np.sum((p - v) ** 2) + t * np.sqrt(np.sum(p * p) + 1e-300)
        best = scipy.optimize.minimize(obj, v, method='Nelder-Mead',
                                       options=dict(xatol=1e-10, fatol=1e-14, maxiter=20000)).x
        got = O.solve_reg(h[None], mu[None], lam, rho)[0]
        assert np.allclose(got, best, atol=2e-4)


def _small_problem(n=24, nt=32, m=4, nz=4, seed=0):
    th = (np.arange(nt) * np.pi / 16).astype(np.float32)
    phi = O.takephi(nt, m)
    rng = np.random.default_rng(seed)
    ftrue = np.zeros((nz, m, n, n))
    yy, xx = np.mgrid[0:n, 0:n]
    for i in range(m):
        ftrue[:, i] = ((xx - 12 - i) ** 2 + (yy - 11) ** 2 < 25)
    ops = O.Operators(n, nt, m, th, phi, n / 2, 4.0)
    g = ops.radonapr(ftrue, 1.0).astype(np.float32)
    return th, phi, g, rng


def test_admm_lambda0_zero_keeps_dual_zero():
    n, nt, m, nzp = 16, 16, 2, 2
    th, phi, g, rng = _small_problem(n=n, nt=nt, m=m, nz=nzp)
    ops = O.Operators(n, nt, m, th, phi, n / 2, 4.0)
    fext = np.zeros((nzp + 2, m, n, n))
    fext[1:-1] = rng.standard_normal((nzp, m, n, n))
    psi = np.zeros((nzp + 1, m + 1, n + 1, n + 1, 4))
    mu = np.zeros_like(psi)
    for _ in range(2):
        O.solver_admm(ops, fext, g, psi, mu, 0.0, 0.5, 0, 0, 2)
        assert np.abs(mu).max() < 1e-12


def test_run_niter_zero_returns_zeros_and_residual_descends():
    n, nt, m, nz = 24, 32, 4, 4
    th, phi, g, _ = _small_problem(n=n, nt=nt, m=m, nz=nz)
    z = O.run(g, th, phi, n, nt, m, nz, 2, n / 2, 0.0, 4.0, 0, 2)
    assert not z.any()
    ops = O.Operators(n, nt, m, th, phi, n / 2, 4.0)
    res = []
    for niter in (1, 3, 6):
        f = O.run(g, th, phi, n, nt, m, nz, 2, n / 2, 1e-5 * 4, 4.0, niter, 2)
        res.append(np.linalg.norm(ops.radonapr(f) - g))
    assert res[0] > res[1] > res[2]
    assert res[2] < 0.5 * np.linalg.norm(g)


def test_check_adjoints_identities_reference_test_shape():
    """The reference test's (tests/test_adjoint.py:14-57) setup at reduced size:
    Fourier basis, theta = linspace(0, 8 pi, ntheta) with endpoint, seed-0 data."""
    n, nt, m, nz = 48, 96, 4, 1
    phi = O.takephi(nt, m)
    np.random.seed(0)
    data = np.random.random([nz, nt, n]).astype('float32')
    theta = np.linspace(0, 8 * np.pi, nt).astype('float32')
    res = O.check_adjoints(data, theta, phi, n, nt, m, nz, nz, 8.0, n / 2)
    assert abs(res[0] - res[1]) < 1e-10 * res[0]
    np.testing.assert_allclose(res[1], res[2], rtol=1)
    np.testing.assert_allclose(res[3], res[4], rtol=1e-1)
    np.testing.assert_allclose(res[4], res[5], rtol=1)
