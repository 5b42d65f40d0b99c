"""Out-of-core solves (state in pinned host memory, windows of chunks streamed
through the device each outer iteration; the reference's managed-memory chunk
loop, rectv.cu:12-16, 143-197): forced with RECTV_STREAM_CHUNKS, the result and
every history column equal the resident solve's (the norms are summed window by
window, a different order of the same FP64 sums), and the trajectory follows the
FP64 oracle at 1e-4 (compared up to the first near-tie of the rho rule)."""
import sys

import numpy as np
import pytest

from oracle import rectv_oracle as O
import oracle_steps as OS

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _problem(n, nz, nrot, ntr, m, seed):
    from rectv_gpu_amd import synthetic as S, basis as B
    th = B.angles(ntr, nrot)
    phi = B.takephi(len(th), m)
    data = S.project(S.spheres(n, nz, 6, seed, nrot=nrot), n, nz, th, ntr, m, noise=0.01, seed=seed)
    return th, phi, data


def _solve(monkeypatch, window, n, nz, nzp, th, phi, data, niter, titer, lam0, lam1):
    from rectv_gpu_amd import Solver
    if window:
        monkeypatch.setenv("RECTV_STREAM_CHUNKS", str(window))
    else:
        monkeypatch.delenv("RECTV_STREAM_CHUNKS", raising=False)
    m = phi.shape[0]
    with Solver(n, len(th), m, nz, nzp, 1) as s:
        s.set_inputs(data, th, phi)
        s.reset()
        s.iterate(niter, titer, lam0 * lam1, lam1)
        return s.result(), s.history(0, niter)


@pytest.mark.parametrize("window", [1, 2])
def test_out_of_core_equals_resident_and_oracle(gpu, monkeypatch, window):
    n, nz, nzp, nrot, ntr, m, niter, titer = 32, 6, 2, 2, 16, 4, 5, 2
    lam0, lam1 = 1e-3, 4.0
    say("case out of core, %d chunk(s) per window" % window, n, nz, nzp, m, niter, titer)
    th, phi, data = _problem(n, nz, nrot, ntr, m, seed=21)
    f_res, h_res = _solve(monkeypatch, 0, n, nz, nzp, th, phi, data, niter, titer, lam0, lam1)
    f_ooc, h_ooc = _solve(monkeypatch, window, n, nz, nzp, th, phi, data, niter, titer, lam0, lam1)
    e = rel(f_ooc, f_res)
    say("  rel L2 out of core vs resident", e)
    assert e < 1e-6
    np.testing.assert_array_equal(h_ooc[:, 2], h_res[:, 2])            # the same rho sequence
    np.testing.assert_allclose(h_ooc, h_res, rtol=1e-9)
    h = []
    ref = O.run(data, th, phi, n, len(th), m, nz, nzp, n / 2, lam0 * lam1, lam1, niter, titer, history=h)
    upto = OS.compare_trajectory(h_ooc, h, rtol=1e-4, say=say)
    if upto == niter:
        assert rel(f_ooc, ref) < 1e-4


def test_out_of_core_without_inner_steps_and_resident_only_calls(gpu, monkeypatch):
    from rectv_gpu_amd import Solver
    n, nz, nzp, nrot, ntr, m = 32, 4, 1, 2, 8, 3
    th, phi, data = _problem(n, nz, nrot, ntr, m, seed=5)
    say("case out of core: titer = 0, resident-only calls")
    f0, h0 = _solve(monkeypatch, 0, n, nz, nzp, th, phi, data, 2, 0, 1e-3, 2.0)
    f1, h1 = _solve(monkeypatch, 3, n, nz, nzp, th, phi, data, 2, 0, 1e-3, 2.0)
    assert not f0.any() and not f1.any()
    np.testing.assert_allclose(h1[:, :5], h0[:, :5], rtol=1e-9)
    monkeypatch.setenv("RECTV_STREAM_CHUNKS", "1")
    with Solver(n, len(th), m, nz, nzp, 1) as s:
        s.set_inputs(data, th, phi)
        with pytest.raises(ValueError, match="out of core"):
            s.get_state()
