"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerances (FP32 arithmetic on the GPU vs FP64 oracle, relative L2):
  * each operator (R_apr, R_apr^*, gradient, divergence): 1e-5
  * check_adjoints values: 1e-5
  * full ADMM reconstruction after niter outer iterations: 1e-4, with the same
    rho sequence.
"""
import sys
import zlib

import numpy as np
import pytest

from oracle import rectv_oracle as O

pytestmark = pytest.mark.gpu

OP_TOL = 1e-5
RUN_TOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _case(name):
    from rectv_gpu_amd import basis as B
    rng = np.random.default_rng(zlib.crc32(name.encode()))  # stable across processes
    if name == "n64_rand":
        n, nt, m, nz = 64, 48, 4, 2
        th = np.sort(rng.uniform(0, 8 * np.pi, nt)).astype(np.float32)
        phi, center = B.takephi(nt, m), n / 2 + 1.25
    elif name == "n72_mixed_radix_m3":
        n, nt, m, nz = 72, 30, 3, 1
        th = B.angles(15, 2)
        phi = rng.standard_normal((m, 2 * nt)).astype(np.float32) / np.sqrt(nt)
        center = n / 2 - 0.5
    elif name == "n384_mixed_radix":      # 2n = 768 = 2^8 * 3: runtime-planned transforms, 8-column workgroups
        n, nt, m, nz = 384, 24, 4, 1
        th = B.angles(12, 2)
        phi, center = B.takephi(nt, m), n / 2 + 0.5
    elif name == "cfg_small":
        n, nt, m, nz = 128, 128, 4, 8
        th = B.angles(32, 4)
        phi, center = B.realbasis(nt, m), n / 2
    elif name == "n64_m20_two_blocks":    # 10 frame pairs: two 16-column MFMA blocks
        n, nt, m, nz = 64, 40, 20, 2
        th = B.angles(20, 2)
        phi, center = B.takephi(nt, m), n / 2 + 0.5
    elif name == "n1024_column_rounds":   # 2n = 2048: two column rounds per workgroup
        n, nt, m, nz = 1024, 12, 3, 1
        th = B.angles(6, 2)
        phi, center = B.takephi(nt, m), n / 2 - 3.5
    elif name == "n2048_column_rounds":   # 2n = 4096: 8-column workgroups
        n, nt, m, nz = 2048, 6, 2, 1
        th = B.angles(3, 2)
        phi, center = B.takephi(nt, m), n / 2
    elif name == "n480_radix5":           # n = 480 = 8*4*3*5, 2n = 960 = 8*8*3*5
        n, nt, m, nz = 480, 10, 3, 1
        th = B.angles(5, 2)
        phi, center = B.takephi(nt, m), n / 2 + 0.5
    elif name == "n504_radix7":           # the reference test's n: 504 = 8*3*3*7, 2n = 1008 = 8*2*3*3*7
        n, nt, m, nz = 504, 12, 4, 1
        th = np.linspace(0, 8 * np.pi, nt).astype(np.float32)
        phi, center = B.takephi(nt, m), n / 2 - 1.5
    elif name == "n40_radix5_small":      # n = 40 = 8*5, 2n = 80 = 16*5
        n, nt, m, nz = 40, 14, 5, 2
        th = B.angles(7, 2)
        phi, center = B.takephi(nt, m), n / 2 + 0.25
    elif name == "n64_m40_pair_groups":   # 20 frame pairs: gather / scatter launches of 16 + 4 pairs
        n, nt, m, nz = 64, 48, 40, 2
        th = B.angles(24, 2)
        phi, center = B.takephi(nt, m), n / 2 + 0.5
    elif name == "n32_m65_pair_groups":   # 33 pairs: 16 + 16 + 1
        n, nt, m, nz = 32, 80, 65, 1
        th = np.sort(rng.uniform(0, 4 * np.pi, nt)).astype(np.float32)
        phi, center = B.takephi(nt, m), n / 2 - 0.5
    elif name == "n36_narrow_tile":       # 2n = 72 = 4*16 + 8: a narrow last tile row / column
        n, nt, m, nz = 36, 20, 4, 2
        th = np.sort(rng.uniform(0, 4 * np.pi, nt)).astype(np.float32)
        phi, center = B.takephi(nt, m), n / 2 + 0.5
    elif name == "n50_three_tiles":       # 2n = 100 = 6*16 + 4: windows meeting three tiles per axis
        n, nt, m, nz = 50, 24, 6, 1
        th = B.angles(12, 2)
        phi, center = B.takephi(nt, m), n / 2 - 0.75
    elif name == "n100_narrow_tile":      # 2n = 200 = 12*16 + 8, m = 18: two column blocks
        n, nt, m, nz = 100, 36, 18, 1
        th = np.sort(rng.uniform(0, 6 * np.pi, nt)).astype(np.float32)
        phi, center = B.takephi(nt, m), n / 2
    elif name == "n44_radix11":           # n = 44 = 4*11, 2n = 88 = 8*11: direct prime stage
        n, nt, m, nz = 44, 18, 4, 1
        th = B.angles(9, 2)
        phi, center = B.takephi(nt, m), n / 2 + 0.5
    elif name == "n34_radix17":           # n = 34 = 2*17, 2n = 68 = 4*17 (a 2448-pixel detector has 17)
        n, nt, m, nz = 34, 16, 3, 2
        th = np.sort(rng.uniform(0, 2 * np.pi, nt)).astype(np.float32)
        phi, center = B.takephi(nt, m), n / 2 - 0.5
    elif name == "n136_radix17":          # n = 136 = 8*17, 2n = 272 = 16*17
        n, nt, m, nz = 136, 12, 2, 1
        th = B.angles(6, 2)
        phi, center = B.takephi(nt, m), n / 2
    elif name == "n45_odd":               # odd n: 2n = 90 = 2 mod 4 (a half-used last column group of H)
        n, nt, m, nz = 45, 20, 4, 2
        th = B.angles(10, 2)
        phi, center = B.takephi(nt, m), n / 2
    elif name == "n17_odd_prime":         # n = 17 (radix 17), 2n = 34: narrow tiles 2 cells wide
        n, nt, m, nz = 17, 12, 3, 1
        th = np.sort(rng.uniform(0, 2 * np.pi, nt)).astype(np.float32)
        phi, center = B.takephi(nt, m), n / 2 + 0.5
    elif name == "n63_odd":               # n = 63 = 9*7, 2n = 126
        n, nt, m, nz = 63, 18, 5, 1
        th = B.angles(9, 2)
        phi, center = B.takephi(nt, m), n / 2 - 1.0
    elif name == "n56_radix7_small":      # n = 56 = 8*7, 2n = 112 = 16*7
        n, nt, m, nz = 56, 15, 3, 2
        th = np.sort(rng.uniform(0, 3 * np.pi, nt)).astype(np.float32)
        phi, center = B.takephi(nt, m), n / 2
    elif name == "n48_m4_packed_partial":  # 2 pairs: 4 slices share the matrix-core columns; 6 = 4 + 2 slices
        n, nt, m, nz = 48, 20, 4, 6
        th = B.angles(10, 2)
        phi, center = B.takephi(nt, m), n / 2 + 0.5
    elif name == "n40_m8_packed_partial":  # 4 pairs: 2 slices share the columns; 3 = 2 + 1 slices
        n, nt, m, nz = 40, 16, 8, 3
        th = np.sort(rng.uniform(0, 4 * np.pi, nt)).astype(np.float32)
        phi, center = B.takephi(nt, m), n / 2 - 0.5
    elif name == "n64_dense_angles":       # 8 rotations x 64 projections: up to 32 items per window
        n, nt, m, nz = 64, 512, 4, 1       # origin (split over blocks), tiles of many blocks and entries
        th = B.angles(64, 8)
        phi, center = B.takephi(nt, m), n / 2 + 0.5
    elif name == "n96_rp3":                # 2n = 192 = 3 * 64: compile-time length 3 * 2^k
        n, nt, m, nz = 96, 24, 5, 3
        th = B.angles(12, 2)
        phi, center = B.takephi(nt, m), n / 2 + 0.5
    elif name == "n160_rp5":               # 2n = 320 = 5 * 64: compile-time length 5 * 2^k
        n, nt, m, nz = 160, 20, 4, 2
        th = np.sort(rng.uniform(0, 4 * np.pi, nt)).astype(np.float32)
        phi, center = B.takephi(nt, m), n / 2 - 0.5
    elif name == "n320_rp5":               # 2n = 640 = 5 * 128
        n, nt, m, nz = 320, 12, 6, 1
        th = B.angles(6, 2)
        phi, center = B.takephi(nt, m), n / 2 + 1.0
    elif name == "n768_rp3":               # 2n = 1536 = 3 * 512 (a 1536-pixel detector doubled)
        n, nt, m, nz = 768, 8, 3, 1
        th = B.angles(4, 2)
        phi, center = B.takephi(nt, m), n / 2 - 2.5
    else:
        raise KeyError(name)
    return n, nt, m, nz, th, phi, center, rng


CASES = ["n64_rand", "n72_mixed_radix_m3", "n384_mixed_radix", "cfg_small", "n64_m20_two_blocks",
         "n1024_column_rounds", "n2048_column_rounds", "n480_radix5", "n504_radix7", "n40_radix5_small",
         "n56_radix7_small", "n64_m40_pair_groups", "n32_m65_pair_groups", "n36_narrow_tile",
         "n50_three_tiles", "n100_narrow_tile", "n44_radix11", "n34_radix17", "n136_radix17", "n45_odd",
         "n17_odd_prime", "n63_odd", "n48_m4_packed_partial", "n40_m8_packed_partial", "n96_rp3",
         "n160_rp5", "n320_rp5", "n768_rp3", "n64_dense_angles"]


@pytest.mark.parametrize("name", CASES)
def test_radon_forward_parity(gpu, name):
    from rectv_gpu_amd import Solver
    n, nt, m, nz, th, phi, center, rng = _case(name)
    say("case", name, "forward R_apr", n, nt, m, nz)
    f = rng.standard_normal((nz, m, n, n)).astype(np.float32)
    g0 = np.zeros((nz, nt, n), np.float32)
    with Solver(n, nt, m, nz, nz, 1) as s:
        s.set_inputs(g0, th, phi, center)
        out = s.apply_radon(f)
    ref = O.Operators(n, nt, m, th, phi, center, 1.0).radonapr(f.astype(np.float64))
    e = rel(out, ref)
    say("  rel L2", e)
    assert e < OP_TOL


@pytest.mark.parametrize("name", CASES)
def test_radon_adjoint_parity(gpu, name):
    from rectv_gpu_amd import Solver
    n, nt, m, nz, th, phi, center, rng = _case(name)
    say("case", name, "adjoint R_apr^*", n, nt, m, nz)
    g = rng.standard_normal((nz, nt, n)).astype(np.float32)
    with Solver(n, nt, m, nz, nz, 1) as s:
        s.set_inputs(g, th, phi, center)
        out = s.apply_radon_adj(g)
    ref = O.Operators(n, nt, m, th, phi, center, 1.0).radonapradj(np.zeros((nz, m, n, n)), g.astype(np.float64))
    e = rel(out, ref)
    say("  rel L2", e)
    assert e < OP_TOL


def test_new_geometry_on_the_same_solver(gpu):
    """The column passes and the scatter touch only the cells of the oversampled
    grid inside some window of the current geometry: after a dense geometry has
    filled the grid, a sparse one set on the same solver still matches the oracle
    in both directions."""
    from rectv_gpu_amd import Solver, basis as B
    n, nt, m, nz = 64, 24, 3, 1
    rng = np.random.default_rng(7)
    dense = np.sort(rng.uniform(0, 4 * np.pi, nt)).astype(np.float32)
    sparse = np.repeat(np.float32([0.3, 1.9]), nt // 2)
    phi = B.takephi(nt, m)
    f = rng.standard_normal((nz, m, n, n)).astype(np.float32)
    g = rng.standard_normal((nz, nt, n)).astype(np.float32)
    say("case geometry change dense -> sparse", n, nt, m)
    with Solver(n, nt, m, nz, nz, 1) as s:
        s.set_inputs(g, dense, phi)
        s.apply_radon(f)
        s.apply_radon_adj(g)
        s.set_inputs(g, sparse, phi)
        fwd = s.apply_radon(f)
        adj = s.apply_radon_adj(g)
    ops = O.Operators(n, nt, m, sparse, phi, n / 2, 1.0)
    e1 = rel(fwd, ops.radonapr(f.astype(np.float64)))
    e2 = rel(adj, ops.radonapradj(np.zeros((nz, m, n, n)), g.astype(np.float64)))
    say("  rel L2 forward", e1, "adjoint", e2)
    assert e1 < OP_TOL and e2 < OP_TOL


def test_gpu_operators_are_adjoint(gpu):
    from rectv_gpu_amd import Solver
    n, nt, m, nz, th, phi, center, rng = _case("n64_rand")
    say("case adjointness of the GPU operators")
    f = rng.standard_normal((nz, m, n, n)).astype(np.float32)
    g = rng.standard_normal((nz, nt, n)).astype(np.float32)
    with Solver(n, nt, m, nz, nz, 1) as s:
        s.set_inputs(g, th, phi, center)
        Rf = s.apply_radon(f)
        Rsg = s.apply_radon_adj(g)
    a = np.vdot(g.astype(np.float64), Rf)
    b = np.vdot(Rsg.astype(np.float64), f)
    say("  <Rf,g>", a, "<f,R*g>", b)
    assert abs(a - b) / abs(a) < 1e-5


@pytest.mark.parametrize("iz", [0, 1, 2])
def test_gradient_divergence_parity(gpu, iz):
    from rectv_gpu_amd import Solver, takephi
    n, nt, m, nzp, nz, lam = 24, 16, 3, 3, 9, 2.5
    say("case gradient/divergence chunk", iz)
    rng = np.random.default_rng(iz)
    fext = rng.standard_normal((nzp + 2, m, n, n)).astype(np.float32)
    flgl, flgr = int(iz != 0), int(iz != nz // nzp - 1)
    with Solver(n, nt, m, nz, nzp, 1) as s:
        s.set_inputs(np.zeros((nz, nt, n), np.float32), np.zeros(nt, np.float32), takephi(nt, m))
        h2 = s.apply_gradient(fext, iz, lam)
        f0 = rng.standard_normal((nzp, m, n, n)).astype(np.float32)
        fd = s.apply_divergence(h2, lam, -0.5, f0)
    ref = O.gradf(O.extendf(fext.astype(np.float64), flgl, flgr), lam)
    e1 = rel(h2, ref)
    refd = O.div(f0.astype(np.float64), ref, -0.5, lam)
    e2 = rel(fd - f0, refd - f0)
    say("  grad", e1, "div", e2)
    assert e1 < OP_TOL and e2 < OP_TOL


def test_check_adjoints_parity(gpu):
    from rectv_gpu_amd import Solver, takephi
    n, nt, m, nz = 64, 64, 4, 1
    say("case check_adjoints", n, nt, m, nz)
    phi = takephi(nt, m)
    np.random.seed(0)
    data = np.random.random([nz, nt, n]).astype('float32')
    theta = np.linspace(0, 8 * np.pi, nt).astype('float32')
    with Solver(n, nt, m, nz, nz, 1) as s:
        res = s.adjoint_tests(data, theta, phi, 8.0)
    ref = O.check_adjoints(data, theta, phi, n, nt, m, nz, nz, 8.0, n / 2)
    say("  gpu", res, "\n  ref", ref)
    assert np.all(np.abs(res - ref) <= OP_TOL * np.abs(ref))


def test_reference_adjoint_test_full_size(gpu):
    """tests/test_adjoint.py:14-57 of the reference at its own size, through the
    reference's import name (import rectv_gpu)."""
    import rectv_gpu
    from rectv_gpu_amd import takephi
    n, ntheta, m, nz, nzp, ngpus, lambda1 = 504, 2400, 16, 1, 1, 1, 8
    say("case reference adjoint test n=504 ntheta=2400 m=16")
    phi = takephi(ntheta, m)
    np.random.seed(0)
    data = np.random.random([nz, ntheta, n]).astype('float32')
    theta = np.linspace(0, 8 * np.pi, ntheta).astype('float32')
    with rectv_gpu.Solver(n, ntheta, m, nz, nzp, ngpus) as slv:
        res = slv.adjoint_tests(data, theta, phi, lambda1)
    say("  res", res)
    np.testing.assert_allclose(res[0], res[1], rtol=1e-1)
    np.testing.assert_allclose(res[1], res[2], rtol=1)
    np.testing.assert_allclose(res[3], res[4], rtol=1e-1)
    np.testing.assert_allclose(res[4], res[5], rtol=1)
    # a true adjoint pair: far tighter than the reference's rtol
    assert abs(res[0] - res[1]) < 1e-4 * res[0]


def _run_case(n, nz, nzp, nrot, ntr, m, niter, titer, seed, basis="fourier", lam0=1e-3, lam1=4.0):
    from rectv_gpu_amd import Solver, synthetic as S, basis as B
    th = B.angles(ntr, nrot)
    nt = len(th)
    phi = B.takephi(nt, m) if basis == "fourier" else B.realbasis(nt, m)
    sph = S.spheres(n, nz, 6, seed, nrot=nrot)
    data = S.project(sph, n, nz, th, ntr, m, noise=0.01, seed=seed)
    with Solver(n, nt, m, nz, nzp, 1) as s:
        s.set_inputs(data, th, phi)
        s.reset()
        s.iterate(niter, titer, lam0 * lam1, lam1)
        hist = s.history(0, niter)
        out = s.result()
    h = []
    ref = O.run(data, th, phi, n, nt, m, nz, nzp, n / 2, lam0 * lam1, lam1, niter, titer, history=h)
    return out, ref, hist, h


@pytest.mark.parametrize("n,nz,nzp,nrot,ntr,m,niter,titer", [
    (32, 6, 2, 2, 16, 4, 5, 2),      # three chunks: halo slabs from the previous iterate
    (64, 4, 4, 2, 32, 4, 4, 3),
    (48, 2, 1, 4, 12, 3, 3, 1),      # nzp = 1, odd m
    (32, 4, 2, 2, 16, 19, 3, 2),     # m = 19: two MFMA column blocks, odd m
    (32, 4, 2, 2, 24, 41, 3, 2),     # m = 41: frame groups 32 + 9, pair groups 16 + 5
    (50, 4, 2, 2, 16, 4, 3, 2),      # n = 50: 2n = 100, narrow last tiles
    (45, 4, 2, 2, 16, 4, 3, 2),      # odd n
    (96, 6, 3, 2, 12, 6, 3, 2),      # 2n = 192 = 3 * 64 (compile-time 3 * 2^k transforms); m = 6: 2 slices packed
    (80, 5, 5, 2, 10, 3, 3, 2),      # 2n = 160: runtime plan; m = 3: 4 slices packed, 5 = 4 + 1
])
def test_run_parity_small(gpu, n, nz, nzp, nrot, ntr, m, niter, titer):
    """Result, rho sequence, r, s, ||f(k+1)-f(k)||^2, ||f(k+1)||^2 and both objective
    terms per outer iteration against the FP64 oracle (trajectory compared up to the
    first near-tie of the rho rule, which is reported)."""
    import oracle_steps as OS
    say("case run", n, nz, nzp, nrot * ntr, m, niter, titer)
    out, ref, hist, h = _run_case(n, nz, nzp, nrot, ntr, m, niter, titer, seed=n + nz)
    e = rel(out, ref)
    say("  rel L2", e, "rho gpu", list(hist[:, 2]), "rho ref", [x["rho"] for x in h])
    upto = OS.compare_trajectory(hist, h, rtol=RUN_TOL, say=say)
    if upto == len(h):
        assert e < RUN_TOL
    lam0 = 1e-3 * 4.0
    J = [x["data"] for x in h[1:]]
    J = np.array(J) + 0.5 * lam0 * np.array([x["tv"] for x in h[:-1]])
    from rectv_gpu_amd import Solver
    say("  objective gpu", Solver.objective(hist, lam0), "oracle", J)
    np.testing.assert_allclose(Solver.objective(hist, lam0), J, rtol=RUN_TOL)


def test_run_in_slice_batches(gpu, monkeypatch):
    """A scratch budget of one chunk forces the solver to run the operators over
    three slice batches (halos, data and frame-pair offsets per batch)."""
    n, nz, nzp, nrot, ntr, m, niter, titer = 32, 6, 2, 2, 16, 4, 3, 2
    say("case run in slice batches", n, nz, nzp, m)
    ref_all = _run_case(n, nz, nzp, nrot, ntr, m, niter, titer, seed=77)
    monkeypatch.setenv("RECTV_SCRATCH_MB", "0.01")
    out, ref, hist, h = _run_case(n, nz, nzp, nrot, ntr, m, niter, titer, seed=77)
    e, e_same = rel(out, ref), rel(out, ref_all[0])
    say("  rel L2 vs oracle", e, "vs one batch", e_same)
    assert e < RUN_TOL and e_same < 1e-6
    # per-batch partial sums of the data term land at their own offsets
    np.testing.assert_allclose(hist[:, 5], ref_all[2][:, 5], rtol=1e-6)
    np.testing.assert_allclose(hist[:, 5], [x["data"] for x in h], rtol=RUN_TOL)


def test_run_parity_small_config(gpu):
    """The plain-method small configuration: N=128, Nz=8, Nrot=4 x 32, M=4 real basis,
    20 spheres, 1% noise, lambda0=1e-3, lambda1=4, niter=32, titer=4."""
    from rectv_gpu_amd import Solver, synthetic as S
    w = S.workload("small")
    say("case run small config", w["n"], w["nz"], w["ntheta"], w["m"], w["niter"])
    titer = 4
    with Solver(w["n"], w["ntheta"], w["m"], w["nz"], w["nzp"], 1) as s:
        out = s.recon(w["data"], w["theta"], w["phi"], lambda0=w["lambda0"], lambda1=w["lambda1"],
                      niter=w["niter"], titer=titer)
    h = []
    ref = O.run(w["data"], w["theta"], w["phi"], w["n"], w["ntheta"], w["m"], w["nz"], w["nzp"], w["center"],
                w["lambda0"] * w["lambda1"], w["lambda1"], w["niter"], titer, history=h)
    e = rel(out, ref)
    say("  rel L2", e)
    assert e < RUN_TOL


def test_run_edge_cases(gpu):
    from rectv_gpu_amd import Solver, takephi
    n, nt, m, nz = 32, 16, 2, 2
    say("case niter=0 / titer=0 / staged == run")
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    phi = takephi(nt, m)
    data = np.random.default_rng(1).random((nz, nt, n)).astype(np.float32)
    with Solver(n, nt, m, nz, 1, 1) as s:
        assert not s.recon(data, th, phi, niter=0).any()
        z = s.recon(data, th, phi, niter=2, titer=0)
        assert np.isfinite(z).all()
        a = s.recon(data, th, phi, lambda0=1e-3, lambda1=2.0, niter=3, titer=2)
        s.set_inputs(data, th, phi)
        s.reset()
        s.iterate(1, 2, 2e-3, 2.0)
        s.iterate(2, 2, 2e-3, 2.0)
        b = s.result()
    assert rel(b, a) < 1e-6
