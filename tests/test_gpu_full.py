"""GPU checks at the full single-GPU workload's own geometry (N=1024, Nz=64, Nzp=16,
Nrot=8 x 200 = 1600 projections, M=16):

  * R_apr and R_apr^* on one slice of that geometry against the FP64 oracle
    (1e-5 relative L2);
  * properties of the full-size solve that hold whatever the data: the first outer
    iteration with one inner step from zero is exactly 1/2 R_apr^* g (the TV and
    dual terms vanish at f = 0; admm.cu:85-97) and its data term is 1/2 ||g||^2;
    <R f, g> = <f, R^* g> over the whole volume; two outer iterations run over four
    slice batches equal the same iterations in one batch.
"""
import sys

import numpy as np
import pytest

from oracle import rectv_oracle as O

pytestmark = pytest.mark.gpu

FULL = dict(n=1024, nrot=8, ntr=200, m=16, nz=64, nzp=16)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.timeout(900)
def test_operators_at_full_geometry_vs_oracle(gpu):
    from rectv_gpu_amd import Solver, basis as B
    n, m = FULL["n"], FULL["m"]
    th = B.angles(FULL["ntr"], FULL["nrot"])
    nt = len(th)
    phi = B.takephi(nt, m)
    rng = np.random.default_rng(1024)
    f = rng.standard_normal((1, m, n, n)).astype(np.float32)
    g = rng.standard_normal((1, nt, n)).astype(np.float32)
    say("case R_apr / R_apr^* at n=1024 ntheta=1600 M=16, one slice")
    with Solver(n, nt, m, 1, 1, 1) as s:
        s.set_inputs(g, th, phi)
        fwd = s.apply_radon(f)
        adj = s.apply_radon_adj(g)
    say("  GPU done; building the FP64 oracle operators")
    ops = O.Operators(n, nt, m, th, phi, n / 2, 1.0)
    e1 = rel(fwd, ops.radonapr(f.astype(np.float64)))
    say("  forward rel L2", e1)
    e2 = rel(adj, ops.radonapradj(np.zeros((1, m, n, n)), g.astype(np.float64)))
    say("  adjoint rel L2", e2)
    assert e1 < 1e-5 and e2 < 1e-5


@pytest.fixture(scope="module")
def full_workload():
    from rectv_gpu_amd import synthetic
    return synthetic.workload("full")


def test_first_iteration_is_half_adjoint_and_adjointness(gpu, full_workload):
    from rectv_gpu_amd import Solver
    w = full_workload
    n, nt, m, nz, nzp = w["n"], w["ntheta"], w["m"], w["nz"], w["nzp"]
    say("case full size: first iteration, adjointness", n, nt, m, nz, nzp)
    with Solver(n, nt, m, nz, nzp, 1) as s:
        s.set_inputs(w["data"], w["theta"], w["phi"])
        s.reset()
        s.iterate(1, 1, 4e-3, 4.0)
        row = s.history(0, 1)[0]
        f1 = s.result()
        rsg = s.apply_radon_adj(w["data"])
        e = rel(f1, 0.5 * rsg)
        g64 = w["data"].astype(np.float64)
        say("  f1 vs R*g/2", e, "data term", row[5], "1/2|g|^2", 0.5 * np.sum(g64 * g64))
        assert e < 1e-6
        assert abs(row[5] - 0.5 * np.sum(g64 * g64)) <= 1e-6 * row[5]
        del f1
        rng = np.random.default_rng(7)
        f = rng.random((nz, m, n, n), dtype=np.float32) - 0.5
        gr = rng.random((nz, nt, n), dtype=np.float32) - 0.5
        rf = s.apply_radon(f)
        a = float(np.dot(rf.ravel().astype(np.float64), gr.ravel().astype(np.float64)))
        del rf
        rg = s.apply_radon_adj(gr)
        b = float(np.dot(rg.ravel().astype(np.float64), f.ravel().astype(np.float64)))
    say("  <Rf,g>", a, "<f,R*g>", b)
    assert abs(a - b) <= 1e-5 * abs(a)


def test_slice_batches_equal_one_batch_at_full_size(gpu, full_workload, monkeypatch):
    from rectv_gpu_amd import Solver
    w = full_workload
    n, nt, m, nz, nzp = w["n"], w["ntheta"], w["m"], w["nz"], w["nzp"]
    lam0, lam1 = w["lambda0"] * w["lambda1"], w["lambda1"]
    out = []
    for budget in (None, "9000"):  # 9000 MiB holds one chunk's scratch (8.2 GB): four batches
        if budget:
            monkeypatch.setenv("RECTV_SCRATCH_MB", budget)
        say("case full size, scratch budget", budget)
        with Solver(n, nt, m, nz, nzp, 1) as s:
            s.set_inputs(w["data"], w["theta"], w["phi"])
            s.reset()
            s.iterate(2, 2, lam0, lam1)
            out.append((s.result(), s.history(0, 2)))
    e = rel(out[1][0], out[0][0])
    say("  four batches vs one", e, out[0][1][:, [0, 1, 2, 5, 6]], out[1][1][:, [0, 1, 2, 5, 6]])
    assert e < 1e-6
    np.testing.assert_allclose(out[1][1], out[0][1], rtol=1e-6)
