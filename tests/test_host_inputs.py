"""Host-side input handling (no GPU): the basis and angle helpers the reference's
test and examples build their inputs with, the array conversions of the drop-in
Solver, and the validation of batch jobs."""
import numpy as np
import pytest

from rectv_gpu_amd import basis as B
from rectv_gpu_amd.solver import _f32
from rectv_gpu_amd.batch import Job


@pytest.mark.parametrize("ntheta,m", [(32, 4), (100, 16), (2400, 16), (15, 3)])
def test_takephi_is_an_orthonormal_fourier_basis(ntheta, m):
    """phi_j[k] = exp(2 pi i a_k b_j / ntheta) / sqrt(ntheta) for integer frequencies
    b_j = j + floor(-m/2): the m rows are orthonormal over the ntheta projections,
    the row for frequency 0 is constant, and rows of opposite frequency are complex
    conjugates."""
    phi = B.takephi(ntheta, m)
    assert phi.shape == (m, 2 * ntheta) and phi.dtype == np.float32
    c = phi[:, ::2].astype(np.float64) + 1j * phi[:, 1::2]
    np.testing.assert_allclose(c @ c.conj().T, np.eye(m), atol=1e-6)
    j0 = (m + 1) // 2  # row of frequency b_j == 0
    np.testing.assert_allclose(c[j0], np.full(ntheta, 1 / np.sqrt(ntheta)), rtol=1e-6)
    if m % 2 == 0 and m >= 4:
        np.testing.assert_allclose(c[j0 + 1], c[j0 - 1].conj(), atol=1e-6)


# The reference builds the basis in its test and examples (tests/test_adjoint.py:22-27,
# examples/foam/foam_example.py:17-23, examples/tao/tao.py:22-28) from index grids
# meshgrid(arange(-ntheta//2, ntheta//2), arange(-m//2, m//2)).  Known answers of those
# grids under Python's floor division, written out by hand:
#   ntheta = 15: k-index -8 .. 6;   ntheta = 16: -8 .. 7
#   m = 3: j-index -2 .. 0;         m = 4: -2 .. 1;   m = 1: -1 .. -1
_REF_GRIDS = {(15, 3): ((-8, 6), (-2, 0)), (16, 3): ((-8, 7), (-2, 0)), (15, 4): ((-8, 6), (-2, 1)),
              (16, 4): ((-8, 7), (-2, 1)), (7, 1): ((-4, 2), (-1, -1))}


@pytest.mark.parametrize("ntheta,m", sorted(_REF_GRIDS))
def test_takephi_equals_reference_basis_for_odd_sizes(ntheta, m):
    """takephi reproduces the reference's basis, including odd ntheta and odd m
    where a grid centred at ntheta//2 would differ (by up to 0.5 per entry)."""
    (k0, k1), (j0, j1) = _REF_GRIDS[(ntheta, m)]
    a = np.arange(k0, k1 + 1, dtype=np.float64)
    b = np.arange(j0, j1 + 1, dtype=np.float64)
    assert len(a) == ntheta and len(b) == m
    ang = 2 * np.pi * np.outer(b, a) / ntheta
    phi = B.takephi(ntheta, m)
    np.testing.assert_allclose(phi[:, ::2], np.cos(ang) / np.sqrt(ntheta), atol=2e-7)
    np.testing.assert_allclose(phi[:, 1::2], np.sin(ang) / np.sqrt(ntheta), atol=2e-7)
    from oracle import rectv_oracle as O
    np.testing.assert_array_equal(O.takephi(ntheta, m), phi)


def test_angles_are_half_turns():
    th = B.angles(200, 8)
    assert th.dtype == np.float32 and len(th) == 1600
    np.testing.assert_allclose(th[200], np.pi, rtol=1e-7)
    np.testing.assert_allclose(np.diff(th.astype(np.float64)), np.pi / 200, rtol=1e-4)


def test_realbasis_is_orthonormal_and_real():
    phi = B.realbasis(128, 4)
    c = phi[:, ::2].astype(np.float64)
    assert not phi[:, 1::2].any()
    np.testing.assert_allclose(c @ c.T, np.eye(4), atol=1e-6)


def test_solver_array_conversion():
    a = np.arange(24, dtype=np.float64).reshape(2, 3, 4)
    out = _f32(a[:, :, ::-1], (2, 3, 4), "data")  # non-contiguous float64
    assert out.dtype == np.float32 and out.flags.c_contiguous
    np.testing.assert_array_equal(out, a[:, :, ::-1])
    flat = _f32(np.zeros(24), (2, 3, 4), "data")  # same size: reshaped
    assert flat.shape == (2, 3, 4)
    with pytest.raises(ValueError):
        _f32(np.zeros(23), (2, 3, 4), "data")


def test_job_validation_and_defaults():
    nt, n, m = 8, 16, 2
    j = Job(np.zeros((3, nt, n)), np.zeros(nt), np.zeros((m, 2 * nt)), niter=4)
    assert (j.nz, j.ntheta, j.n, j.m, j.nzp) == (3, nt, n, m, 3)
    assert j.rot_center == n / 2 and j.voxels == n * n * 3 * m
    with pytest.raises(ValueError):
        Job(np.zeros((nt, n)), np.zeros(nt), np.zeros((m, 2 * nt)))
    with pytest.raises(ValueError):
        Job(np.zeros((1, nt, n)), np.zeros(nt), np.zeros((m, nt)))
    k = j.fresh()
    assert k is not j and k.data is j.data and k.result is None and k.name == j.name
