"""GPU parity of the fused ADMM kernels one phase at a time, through the C ABI:

  * one inner step (k_ray's residual, the scatter / inverse transforms and
    k_update's fused TV + R^* update, with div(mu - rho psi) from k_divchi) against
    admm.cu:85-97 composed from the FP64 oracle operators;
  * the closing phase (k_prox -> k_reduce -> k_rho_rule -> k_divchi) against
    admm.cu:101-117 and rectv.cu:224-232: psi, mu, div(mu - rho' psi) and each of
    r, s, ||fn - f||^2, ||fn||^2, sum |grad fn| individually, and the new rho;
  * the state round trip and a resumed run (checkpoint / restore).

Tolerance: 1e-5 relative L2 per output (FP32 kernels against FP64), the update
fout - fcur compared on its own so that the unchanged part of fcur cannot hide an
error in it.
"""
import sys

import numpy as np
import pytest

from oracle import rectv_oracle as O
import oracle_steps as OS

pytestmark = pytest.mark.gpu

TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _state(seed, n, nt, m, nz, nzp):
    from rectv_gpu_amd import basis as B
    rng = np.random.default_rng(seed)
    th = np.sort(rng.uniform(0, 4 * np.pi, nt)).astype(np.float32)
    phi = B.takephi(nt, m)
    nch = nz // nzp
    A = rng.standard_normal((nz, m, n, n)).astype(np.float32)
    fcur = (A + 0.3 * rng.standard_normal(A.shape)).astype(np.float32)
    shp = (nch, nzp + 1, m + 1, n + 1, n + 1, 4)
    psi = (0.2 * rng.standard_normal(shp)).astype(np.float32)
    mu = (0.1 * rng.standard_normal(shp)).astype(np.float32)
    g = rng.random((nz, nt, n)).astype(np.float32)
    return th, phi, A, fcur, psi, mu, g


# n=32: 2n = 64 compile-time transforms; n=40: 2n = 80 = 16 * 5 runtime-planned path;
# m odd (a half-empty last frame pair); three chunks (halo slabs both sides);
# m = 67: frame groups of 32 + 32 + 3 in the TV kernels, pair groups of 16 + 16 + 2
CASES = [(32, 24, 5, 6, 2, 0.75, 2.5), (40, 20, 4, 4, 2, 2.0, 4.0), (64, 16, 3, 3, 1, 0.25, 1.0),
         (32, 72, 67, 4, 2, 1.5, 3.0), (36, 20, 4, 4, 2, 0.5, 2.0), (45, 16, 3, 4, 2, 1.0, 2.0)]


@pytest.mark.parametrize("n,nt,m,nz,nzp,rho,lam1", CASES)
def test_inner_step_parity(gpu, n, nt, m, nz, nzp, rho, lam1):
    from rectv_gpu_amd import Solver
    say("case inner step", n, nt, m, nz, nzp, rho, lam1)
    th, phi, A, fcur, psi, mu, g = _state(n + m, n, nt, m, nz, nzp)
    with Solver(n, nt, m, nz, nzp, 1) as s:
        s.set_inputs(g, th, phi)
        s.set_state(A, psi, mu, rho, lam1)
        out = s.apply_inner_step(fcur, lam1)
    ops = O.Operators(n, nt, m, th, phi, n / 2, lam1)
    ref = OS.inner_step(ops, A.astype(np.float64), fcur.astype(np.float64), g.astype(np.float64),
                        psi.astype(np.float64), mu.astype(np.float64), rho, nzp)
    # the two parts of the update, for scale
    tv = OS.inner_step(ops, A.astype(np.float64), fcur.astype(np.float64),
                       ops.radonapr(fcur.astype(np.float64)), psi.astype(np.float64), mu.astype(np.float64), rho, nzp)
    d_ref = ref - fcur
    e = rel(out - fcur, d_ref)
    say("  update rel L2", e, "| |TV part| %.3g |data part| %.3g" % (np.linalg.norm(tv - fcur),
                                                                    np.linalg.norm(ref - tv)))
    assert e < TOL


@pytest.mark.parametrize("n,nt,m,nz,nzp,rho,lam1", CASES)
def test_prox_phase_parity(gpu, n, nt, m, nz, nzp, rho, lam1):
    from rectv_gpu_amd import Solver
    th, phi, A, fcur, psi, mu, g = _state(2 * n + m, n, nt, m, nz, nzp)
    ops = O.Operators(n, nt, m, th, phi, n / 2, lam1)
    # threshold at the median group norm of grad fcur + mu/rho: both branches of the shrink
    fe, fl, fr = OS.chunk_ext(fcur.astype(np.float64), A, 0, nzp, nz // nzp)
    v = ops.gradient(fe, fl, fr) + mu[0] / rho
    lam0 = float(rho * np.median(np.linalg.norm(v, axis=-1)))
    say("case closing phase", n, nt, m, nz, nzp, rho, lam0, lam1)
    with Solver(n, nt, m, nz, nzp, 1) as s:
        s.set_inputs(g, th, phi)
        s.set_state(A, psi, mu, rho, lam1)
        row = s.apply_prox(fcur, lam0, lam1)
        st = s.get_state()
    ref = OS.prox_step(ops, A.astype(np.float64), fcur.astype(np.float64), psi.astype(np.float64),
                       mu.astype(np.float64), rho, lam0, lam1, nzp)
    shrunk = np.mean(np.all(ref["psi"] == 0, axis=-1))
    errs = {k: rel(st[k], ref[k]) for k in ("psi", "mu", "dchi")}
    errs.update({k: abs(row[c] - ref[k]) / abs(ref[k]) for c, k in ((0, "r"), (1, "s"), (3, "rdiff"),
                                                                    (4, "fnorm"), (6, "tv"))})
    say("  ", {k: "%.2e" % v for k, v in errs.items()}, "groups shrunk to 0: %.2f" % shrunk,
        "rho", row[2], "->", st["rho"], "(oracle", ref["rho_next"], ")")
    assert 0.05 < shrunk < 0.95
    assert all(v < TOL for v in errs.values()), errs
    assert row[2] == rho and st["rho"] == ref["rho_next"]
    assert np.isnan(row[5])  # no inner step ran: no data term
    np.testing.assert_array_equal(st["f"], A)  # the previous iterate is kept


def test_prox_phase_parity_rho_halved(gpu):
    """The closing phase when the rule halves rho (s > 10 r: a small threshold leaves
    psi close to grad fcur + mu/rho while fcur moved far from the previous iterate):
    div(mu - rho' psi) is then formed from chi and mu with rho' = rho / 2."""
    from rectv_gpu_amd import Solver
    n, nt, m, nz, nzp, rho, lam1 = 40, 20, 4, 4, 2, 2.0, 1.0
    th, phi, A, fcur, psi, mu, g = _state(7, n, nt, m, nz, nzp)
    ops = O.Operators(n, nt, m, th, phi, n / 2, lam1)
    fe, fl, fr = OS.chunk_ext(fcur.astype(np.float64), A, 0, nzp, nz // nzp)
    v = ops.gradient(fe, fl, fr) + mu[0] / rho
    lam0 = float(1e-3 * rho * np.median(np.linalg.norm(v, axis=-1)))
    say("case closing phase, rho halved", n, nt, m, nz, nzp, rho, lam0, lam1)
    with Solver(n, nt, m, nz, nzp, 1) as s:
        s.set_inputs(g, th, phi)
        s.set_state(A, psi, mu, rho, lam1)
        row = s.apply_prox(fcur, lam0, lam1)
        st = s.get_state()
    ref = OS.prox_step(ops, A.astype(np.float64), fcur.astype(np.float64), psi.astype(np.float64),
                       mu.astype(np.float64), rho, lam0, lam1, nzp)
    assert ref["rho_next"] == rho / 2, ref["rho_next"]
    errs = {k: rel(st[k], ref[k]) for k in ("psi", "dchi")}
    # mu' = rho (p - psi) with p = grad fcur + mu/rho is a small difference here
    # (|mu'| <= lambda0 per group, far below |mu| + rho |grad fcur|), so its error is
    # measured against the scale of the dual it was formed from
    errs["mu"] = float(np.linalg.norm(st["mu"] - ref["mu"]) / np.linalg.norm(mu))
    errs.update({k: abs(row[c] - ref[k]) / abs(ref[k]) for c, k in ((0, "r"), (1, "s"))})
    say("  ", {k: "%.2e" % v for k, v in errs.items()}, "rho", row[2], "->", st["rho"])
    assert all(v < TOL for v in errs.values()), errs
    assert st["rho"] == rho / 2


def test_state_round_trip_and_resume(gpu):
    """A run of 2 + 3 outer iterations with the state saved after 2 and restored into
    a fresh solver equals 5 iterations in one go (1e-6).  The solver keeps
    chi = mu - rho psi in place of psi, so the exported psi = (mu - chi) / rho and the
    div(mu - rho psi) re-formed from it round-trip to FP32 rounding (1e-6), f and mu
    bitwise."""
    from rectv_gpu_amd import Solver, synthetic as S, basis as B
    n, nz, nzp, nrot, ntr, m, titer = 32, 6, 2, 2, 16, 4, 2
    th = B.angles(ntr, nrot)
    nt = len(th)
    phi = B.takephi(nt, m)
    data = S.project(S.spheres(n, nz, 6, 5, nrot=nrot), n, nz, th, ntr, m, noise=0.01, seed=5)
    lam0, lam1 = 4e-3, 4.0
    say("case checkpoint / resume", n, nz, nzp, nt, m)
    with Solver(n, nt, m, nz, nzp, 1) as s:
        s.set_inputs(data, th, phi)
        s.reset()
        h5 = s.iterate(5, titer, lam0, lam1, history=True)
        full = s.result()
    with Solver(n, nt, m, nz, nzp, 1) as s:
        s.set_inputs(data, th, phi)
        s.reset()
        h2 = s.iterate(2, titer, lam0, lam1, history=True)
        st = s.get_state()
    np.testing.assert_array_equal(h2[:, 2], h5[:2, 2])
    with Solver(n, nt, m, nz, nzp, 1) as s:
        s.set_inputs(data, th, phi)
        s.set_state(st["f"], st["psi"], st["mu"], st["rho"], lam1)
        st2 = s.get_state()
        for k in ("f", "mu"):
            np.testing.assert_array_equal(st2[k], st[k])
        for k in ("psi", "dchi"):
            say("  round trip", k, rel(st2[k], st[k]))
            assert rel(st2[k], st[k]) < 1e-6
        h3 = s.iterate(3, titer, lam0, lam1, history=True)
        resumed = s.result()
    say("  resumed vs one run", rel(resumed, full))
    np.testing.assert_array_equal(h3[:, 2], h5[2:, 2])
    assert rel(resumed, full) < 1e-6
