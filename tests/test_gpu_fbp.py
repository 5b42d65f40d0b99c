"""GPU parity of the filtered back-projection, the support mask and the warm start
(the reference's unused applyfilter / circ kernels, kernels_radonusfft.cuh:146-189,
and the `filter` argument of radonusfft::adjR, radonusfft.cu:106) against the FP64
oracle: the FBP at 1e-5 relative L2 with and without the mask (the mask's zero
pattern exactly); masked and warm-started runs at 1e-4 with their trajectories;
the warm start lowers the objective reached after a few iterations."""
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


def _phantom_data(n, nz, nrot, ntr, m, seed):
    from rectv_gpu_amd import synthetic as S, basis as B
    th = B.angles(ntr, nrot)
    phi = B.takephi(len(th), m)
    data = S.project(S.spheres(n, nz, 6, seed, nrot=nrot), n, nz, th, ntr, m, noise=0.01, seed=seed)
    return th, phi, data


@pytest.mark.parametrize("n,nz,nrot,ntr,m,r", [(64, 2, 2, 32, 4, None), (64, 2, 2, 32, 4, 0.05),
                                               (72, 1, 2, 15, 3, 0.0), (128, 2, 4, 32, 16, 0.2)])
def test_fbp_parity(gpu, n, nz, nrot, ntr, m, r):
    from rectv_gpu_amd import Solver
    th, phi, data = _phantom_data(n, nz, nrot, ntr, m, seed=n + m)
    nt = len(th)
    say("case FBP", n, nz, nt, m, "support", r)
    with Solver(n, nt, m, nz, nz, 1) as s:
        s.set_support(r)
        out = s.fbp(data, th, phi)
    ref = O.Operators(n, nt, m, th, phi, n / 2, 1.0).fbp(data.astype(np.float64), r=r)
    e = rel(out, ref)
    say("  rel L2", e)
    assert e < 1e-5
    if r is not None:
        np.testing.assert_array_equal(out[0, 0] != 0, O.circ_mask(n, r) != 0)


def _run_pair(n, nz, nzp, nrot, ntr, m, niter, titer, support, init, seed):
    from rectv_gpu_amd import Solver
    th, phi, data = _phantom_data(n, nz, nrot, ntr, m, seed)
    nt = len(th)
    lam0, lam1 = 1e-3, 4.0
    with Solver(n, nt, m, nz, nzp, 1) as s:
        out = s.recon(data, th, phi, lambda0=lam0, lambda1=lam1, niter=niter, titer=titer, init=init,
                      support=support)
        hist = s.history(0, niter)
    ops = O.Operators(n, nt, m, th, phi, n / 2, lam1)
    state = {"f": ops.fbp(data.astype(np.float64), r=support)} if init == "fbp" else None
    h = []
    ref = O.run(data, th, phi, n, nt, m, nz, nzp, n / 2, lam0 * lam1, lam1, niter, titer, history=h,
                state=state, support=support)
    return out, ref, hist, h


@pytest.mark.parametrize("init,support", [(None, 0.1), ("fbp", None), ("fbp", 0.05)])
def test_masked_and_warm_started_runs(gpu, init, support):
    n, nz, nzp, nrot, ntr, m, niter, titer = 64, 4, 2, 2, 32, 4, 4, 2
    say("case run init", init, "support", support)
    out, ref, hist, h = _run_pair(n, nz, nzp, nrot, ntr, m, niter, titer, support, init, seed=9)
    e = rel(out, ref)
    say("  rel L2", e)
    upto = OS.compare_trajectory(hist, h, rtol=1e-4, say=say)
    if upto == niter:
        assert e < 1e-4
    if support is not None:
        assert not out[..., O.circ_mask(n, support) == 0].any()


def test_warm_start_lowers_the_objective(gpu):
    from rectv_gpu_amd import Solver
    n, nz, nzp, nrot, ntr, m, titer = 64, 4, 4, 2, 32, 4, 2
    th, phi, data = _phantom_data(n, nz, nrot, ntr, m, seed=3)
    nt = len(th)
    lam0 = 4e-3
    J = {}
    for init in ("zero", "fbp"):
        with Solver(n, nt, m, nz, nzp, 1) as s:
            s.recon(data, th, phi, lambda0=1e-3, lambda1=4.0, niter=4, titer=titer, init=init, support=0.05)
            J[init] = Solver.objective(s.history(0, 4), lam0)
    say("  objective from zero", J["zero"], "from FBP", J["fbp"])
    assert np.all(J["fbp"] < J["zero"])
