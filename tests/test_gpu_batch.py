"""GPU: the asynchronous, graph-replayed, batched and served forms give the same
reconstructions as a plain solve of each job on its own, and the served
results stay within the FP64-oracle tolerance.

Tolerances: graph replay, async enqueue, batch and server runs execute the same
kernels in the same order on each job's own buffers, so they must agree with the
plain solve to 1e-6 relative L2 (they are expected to be bitwise equal); a served
job against the FP64 oracle: 1e-4 relative L2 (the full-solve tolerance).
"""
import sys
import threading
import time

import numpy as np
import pytest

from oracle import rectv_oracle as O

pytestmark = pytest.mark.gpu

SAME_TOL = 1e-6
RUN_TOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _job(n, nz, nzp, nrot, ntr, m, niter, titer, seed, tol=None):
    from rectv_gpu_amd import synthetic as S, basis as B
    from rectv_gpu_amd.batch import Job
    th = B.angles(ntr, nrot)
    nt = len(th)
    phi = B.takephi(nt, m)
    sph = S.spheres(n, nz, 6, seed, nrot=nrot)
    data = S.project(sph, n, nz, th, ntr, m, noise=0.01, seed=seed)
    return Job(data, th, phi, lambda0=1e-3, lambda1=4.0, niter=niter, titer=titer, nzp=nzp, tol=tol,
               name="n%d_nz%d_m%d_s%d" % (n, nz, m, seed))


def _alone(j):
    from rectv_gpu_amd import Solver
    with Solver(j.n, j.ntheta, j.m, j.nz, j.nzp, 1) as s:
        return s.recon(j.data, j.theta, j.phi, j.rot_center, j.lambda0, j.lambda1, j.niter, j.titer)


def test_graph_and_async_match_eager(gpu):
    from rectv_gpu_amd import Solver
    j = _job(64, 4, 2, 2, 32, 4, 5, 2, seed=3)
    say("case graph replay / async enqueue vs eager", j.name)
    with Solver(j.n, j.ntheta, j.m, j.nz, j.nzp, 1) as s:
        s.set_inputs(j.data, j.theta, j.phi)
        s.reset()
        h0 = s.iterate(j.niter, j.titer, 4e-3, 4.0, history=True)
        a = s.result()
        s.reset()
        s.set_graph(True)
        h1 = s.iterate(j.niter, j.titer, 4e-3, 4.0, history=True)
        b = s.result()
        s.reset()
        s.iterate_async(2, j.titer, 4e-3, 4.0)
        s.iterate_async(3, j.titer, 4e-3, 4.0)
        while not s.idle():
            time.sleep(1e-3)
        assert s.iterations() == 5
        h2 = s.history(0, 5)
        c = s.result()
    say("  graph", rel(b, a), "async", rel(c, a))
    assert rel(b, a) < SAME_TOL and rel(c, a) < SAME_TOL
    np.testing.assert_array_equal(h0[:, 2], h1[:, 2])
    np.testing.assert_array_equal(h0[:, 2], h2[:, 2])
    np.testing.assert_allclose(h0[:, 3], h2[:, 3], rtol=1e-6)
    assert np.all(h2[:, 4] > 0)


def test_autotune_keeps_results(gpu):
    from rectv_gpu_amd import Solver
    j = _job(64, 8, 4, 2, 32, 6, 2, 2, seed=5)
    say("case autotune", j.name)
    a = _alone(j)
    with Solver(j.n, j.ntheta, j.m, j.nz, j.nzp, 1) as s:
        s.set_inputs(j.data, j.theta, j.phi)
        s.reset()
        zg, zw = s.autotune()
        say("  choice gather slices", zg, "scatter slices", zw)
        assert zg in (1, 2, 4, 8, 16, 32, 64) and zw in (2, 4, 8)
        s.reset()
        s.iterate(j.niter, j.titer, j.lambda0 * j.lambda1, j.lambda1)
        b = s.result()
    assert rel(b, a) < SAME_TOL


def test_footprint_bounds_device_bytes(gpu):
    from rectv_gpu_amd import Solver
    for (n, nz, nzp, nrot, ntr, m) in [(64, 4, 2, 2, 32, 4), (128, 2, 2, 4, 16, 8), (96, 3, 1, 2, 20, 19)]:
        j = _job(n, nz, nzp, nrot, ntr, m, 1, 1, seed=n + m)
        fp = Solver.footprint(j.n, j.ntheta, j.m, j.nz, j.nzp)
        with Solver(j.n, j.ntheta, j.m, j.nz, j.nzp, 1) as s:
            s.set_inputs(j.data, j.theta, j.phi)
            used = s.device_bytes()
        say("case footprint", j.name, "bound", fp, "held", used)
        assert 0 < used <= fp


def test_batch_matches_alone(gpu):
    from rectv_gpu_amd.batch import solve_batch
    jobs = [_job(64, 4, 2, 2, 32, 4, 4, 2, seed=11), _job(128, 2, 2, 4, 16, 8, 3, 1, seed=12),
            _job(32, 6, 3, 2, 16, 3, 5, 2, seed=13)]
    say("case batch of", [j.name for j in jobs])
    ref = [_alone(j) for j in jobs]
    solve_batch(jobs, workers=3)
    for j, r in zip(jobs, ref):
        e = rel(j.result, r)
        say("  ", j.name, e, "iterations", j.iterations)
        assert j.error is None and j.iterations == j.niter
        assert e < SAME_TOL
    # the C round robin over prepared solvers (rtv_batch_iterate)
    from rectv_gpu_amd.batch import prepare, batch_iterate
    sol = [prepare(j, autotune=False) for j in jobs]
    try:
        batch_iterate(sol, [j.niter for j in jobs], [j.titer for j in jobs],
                      [j.lambda0 * j.lambda1 for j in jobs], [j.lambda1 for j in jobs])
        for s, j, r in zip(sol, jobs, ref):
            assert s.iterations() == j.niter
            e = rel(s.result(), r)
            say("  round robin", j.name, e)
            assert e < SAME_TOL
    finally:
        for s in sol:
            s.free()


def test_server_mix_matches_alone_and_oracle(gpu):
    from rectv_gpu_amd.batch import Server
    jobs = [_job(64, 4, 2, 2, 32, 4, 4, 2, seed=21), _job(32, 2, 1, 2, 16, 3, 6, 1, seed=22),
            _job(128, 2, 2, 4, 16, 8, 3, 2, seed=23), _job(64, 2, 2, 2, 24, 4, 5, 2, seed=24),
            _job(32, 4, 4, 2, 16, 2, 40, 2, seed=25, tol=5e-3)]
    say("case server mix of", [j.name for j in jobs])
    ref = [_alone(j) for j in jobs[:4]]
    srv = Server(max_active=3, ahead=1, workers=2)
    try:
        # the early-stopping job and two others arrive together, the rest while they run
        order = [jobs[4], jobs[0], jobs[1], jobs[2], jobs[3]]

        def producer():
            for q, j in enumerate(order):
                srv.submit(j)
                if q >= 2:
                    time.sleep(0.02)
        th = threading.Thread(target=producer)
        th.start()
        srv.run(until=lambda: not th.is_alive() and srv.pending() == 0)
        th.join()
    finally:
        srv.close()
    for j in jobs:
        assert j.error is None, j.error
        assert j.done.is_set() and j.latency() > 0
    for j, r in zip(jobs[:4], ref):
        e = rel(j.result, r)
        say("  ", j.name, "mix vs alone", e)
        assert j.iterations == j.niter and e < SAME_TOL
    # the early-stopped job: stopped at the first iteration whose relative change is below tol
    js = jobs[4]
    rc = np.sqrt(js.history[:, 3] / js.history[:, 4])
    say("  tol job stopped after", js.iterations, "relative changes", rc)
    assert js.iterations < js.niter and rc[-1] < js.tol and np.all(rc[:-1] >= js.tol)
    # the served results against the FP64 oracle (one job of the mix)
    j = jobs[1]
    ref64 = O.run(j.data, j.theta, j.phi, j.n, j.ntheta, j.m, j.nz, j.nzp, j.rot_center,
                  j.lambda0 * j.lambda1, j.lambda1, j.niter, j.titer)
    e = rel(j.result, ref64)
    say("  served job vs FP64 oracle", e)
    assert e < RUN_TOL
    assert srv.peak_active >= 2
