"""Pins for the solve level of the FP64 oracle: the basis orientation, the frame
times, the chunk loop, the rho rule and the ADMM fixed point.  Each answer is
fixed by the mathematics or by the reference's integer arithmetic, not by the
oracle's own formulas:

  * basis orientation: with M = ntheta the takephi basis is a full DFT, so
    sum_j phi_j[k] conj(phi_j[i]) = delta_ki and row k of R_apr f is
    sqrt(M) Re fwdR(f_k) row k (frame k seen at angle k; operators.cu:22-31,
    kernels_operators.cuh:99-149); the adjoint sends row i of g to frame i only;
  * frame times: t_i = (int)(i*ntheta/m) (kernels_operators.cuh:113, :130) in C
    integer arithmetic, known answers written out by hand;
  * chunk loop: chunks read the previous iterate f and write fn (rectv.cu:170-171),
    so the order in which chunks are processed cannot change anything;
  * rho rule: residual balancing (Boyd et al. 2011, Distributed Optimization and
    Statistical Learning via ADMM, sec. 3.4.1; rectv.cu:224-232): with the primal
    residual r = ||grad f - psi||^2 and the dual residual s = rho^2 ||grad(f_old -
    f_new)||^2 computed here from the iterates, r >> s raises rho, s >> r lowers it;
  * ADMM fixed point: a converged run satisfies the stationarity conditions of
    its own update rules (admm.cu:85-111): psi = grad f, mu = lambda0 psi/|psi|
    on groups with psi != 0 and |mu| <= lambda0 elsewhere, and
    R^*(R f - g) + div(mu) = 0 (the inner step's fixed point); the objective
    1/2||R f - g||^2 + lambda0/2 sum|grad f| settles at its late minimum.

The pins are functions of the oracle module so that tests/test_oracle_mutations.py
can run them against copies of the oracle with planted mistakes.
"""
import numpy as np

from oracle import rectv_oracle as O_REAL


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def _angles(nt, seed, span):
    rng = np.random.default_rng(seed)
    return np.sort(rng.uniform(0, span, nt)).astype(np.float32)


# ---------------------------------------------------------------- pins
def pin_basis_orientation(O):
    n, M = 16, 8
    nt = M
    th = _angles(nt, 21, 4 * np.pi)
    center = n / 2 + 0.5
    phi = O.takephi(nt, M)
    ops = O.Operators(n, nt, M, th, phi, center, 1.0)
    U = O.Usfft(n, th, center)
    rng = np.random.default_rng(22)
    f = rng.standard_normal((2, M, n, n))
    Rf = ops.radonapr(f, 1.0)
    for k in range(nt):
        expect = np.sqrt(M) * U.fwdR(f[:, k].astype(complex))[:, k, :].real
        assert _rel(Rf[:, k], expect) < 1e-5, ("forward row", k)
    g = rng.standard_normal((2, nt, n))
    Rsg = ops.radonapradj(np.zeros_like(f), g, 1.0)
    for i in range(M):
        gi = np.zeros_like(g)
        gi[:, i] = g[:, i]
        expect = np.sqrt(M) * U.adjR(gi.astype(complex)).real
        assert _rel(Rsg[:, i], expect) < 1e-5, ("adjoint frame", i)


# (int)(i*ntheta/m) for i = 0..m-1, evaluated by hand in C integer arithmetic
_FRAME_TIMES = {
    (200, 16): [0, 12, 25, 37, 50, 62, 75, 87, 100, 112, 125, 137, 150, 162, 175, 187],
    (15, 4): [0, 3, 7, 11],
    (100, 8): [0, 12, 25, 37, 50, 62, 75, 87],
    (7, 3): [0, 2, 4],
}


def pin_frame_times(O):
    """decphi (u = sum_i f_i conj phi[t_i]) and recphi (f_i += u phi[t_i]) probed
    with the ramp basis phi[k] = k: frame i reads and writes basis sample t_i."""
    for (nt, m), want in _FRAME_TIMES.items():
        ramp = np.arange(nt, dtype=np.float64) + 0j
        got = []
        for i in range(m):
            fc = np.zeros((1, m, 1, 1), complex)
            fc[0, i] = 1.0
            got.append(int(round(O.decphi(fc, ramp, nt, m).real[0, 0, 0])))
        assert got == want, ("decphi", nt, m, got)
        acc = np.zeros((1, m, 1, 1), complex)
        O.recphi(acc, np.ones((1, 1, 1)), ramp, nt, m)
        assert [int(round(v)) for v in acc.real[0, :, 0, 0]] == want, ("recphi", nt, m)


def pin_chunk_order(O):
    n, nt, m, nz, nzp = 16, 8, 2, 6, 2
    th = _angles(nt, 23, 2 * np.pi)
    phi = O.takephi(nt, m)
    g = np.random.default_rng(24).random((nz, nt, n)).astype(np.float32)
    args = (g, th, phi, n, nt, m, nz, nzp, n / 2, 1e-2, 4.0, 3, 2)
    h0, h1, h2 = [], [], []
    a = O.run(*args, history=h0)
    b = O.run(*args, history=h1, order=[2, 1, 0])
    c = O.run(*args, history=h2, order=[1, 2, 0])
    d = O.run(*args, threads=3)
    assert _rel(b, a) < 1e-12 and _rel(c, a) < 1e-12 and _rel(d, a) < 1e-12
    for x, y in zip(h0, h1):
        assert x["rho"] == y["rho"] and abs(x["r"] - y["r"]) <= 1e-12 * abs(x["r"])


def _grad(O, ops, f):
    return ops.gradient(np.concatenate([f[:1], f, f[-1:]], axis=0), 0, 0)


def pin_rho_direction(O):
    n, nt, m, nz = 16, 8, 2, 2
    th = _angles(nt, 25, 2 * np.pi)
    phi = O.takephi(nt, m)
    rng = np.random.default_rng(26)
    g = rng.random((nz, nt, n)).astype(np.float32)
    ops = O.Operators(n, nt, m, th, phi, n / 2, 4.0)
    yy, xx = np.mgrid[0:n, 0:n]
    f0 = np.broadcast_to(np.sin(xx / 3.0) * np.cos(yy / 4.0), (nz, m, n, n)).copy()
    cases = [
        # a huge penalty weight keeps psi = 0 and titer = 0 keeps f: r = |grad f|^2, s = 0
        dict(lam0=1e6, titer=0, start=f0, expect="up"),
        # no penalty weight: psi = grad f exactly, so r = 0, while f moves: s > 0
        dict(lam0=0.0, titer=1, start=np.zeros_like(f0), expect="down"),
    ]
    for cs in cases:
        st = {"f": cs["start"], "rho": 0.5}
        h = []
        O.run(g, th, phi, n, nt, m, nz, nz, n / 2, cs["lam0"], 4.0, 1, cs["titer"], history=h, state=st)
        gnew = _grad(O, ops, st["f"])
        r = float(np.sum((gnew - st["psi"][0]) ** 2))
        s = 0.25 * float(np.sum((_grad(O, ops, cs["start"]) - gnew) ** 2))
        # the history reports the same residuals
        assert abs(h[0]["r"] - r) <= 1e-9 * max(r, 1e-30) + 1e-30
        assert abs(h[0]["s"] - s) <= 1e-9 * max(s, 1e-30) + 1e-30
        if cs["expect"] == "up":
            assert r > 10 * s and h[0]["rho_next"] > h[0]["rho"], (r, s, h[0])
        else:
            assert s > 10 * r and h[0]["rho_next"] < h[0]["rho"], (r, s, h[0])


def pin_admm_fixed_point(O, niter=200):
    n, nt, m, nz = 16, 16, 2, 2
    lam0, lam1 = 0.05, 4.0
    th = (np.arange(nt) * np.pi / 8).astype(np.float32)
    phi = O.takephi(nt, m)
    ops = O.Operators(n, nt, m, th, phi, n / 2, lam1)
    yy, xx = np.mgrid[0:n, 0:n]
    ftrue = np.zeros((nz, m, n, n))
    for i in range(m):
        ftrue[:, i] = ((xx - 8 - i) ** 2 + (yy - 7) ** 2 < 12) * 1.0
    g = (ops.radonapr(ftrue) + 0.02 * np.random.default_rng(0).standard_normal((nz, nt, n))).astype(np.float32)
    st, h = {}, []
    f = O.run(g, th, phi, n, nt, m, nz, nz, n / 2, lam0, lam1, niter, 2, history=h, state=st)
    assert np.isfinite(f).all()
    G = _grad(O, ops, f)
    psi, mu = st["psi"][0], st["mu"][0]
    # primal feasibility
    assert _rel(psi, G) < 2e-3
    # dual variables in the group subdifferential of lam0 |psi|
    npsi = np.linalg.norm(psi, axis=-1)
    on = npsi > 1e-6 * npsi.max()
    assert 0.1 < on.mean() < 0.9  # both kinds of groups occur
    assert np.abs(mu[on] - lam0 * psi[on] / npsi[on][:, None]).max() < 1e-6 * lam0
    assert np.linalg.norm(mu[~on], axis=-1).max() <= lam0 * (1 + 1e-6)
    # stationarity of the inner step: R^*(R f - g) + div mu = 0
    g64 = g.astype(np.float64)
    grad_data = ops.radonapradj(np.zeros_like(f), ops.radonapr(f) - g64)
    dm = O.div(np.zeros_like(f), mu, 1.0, lam1)
    scale = np.linalg.norm(ops.radonapradj(np.zeros_like(f), g64))
    assert np.linalg.norm(grad_data + dm) < 2e-4 * scale
    # the objective settles without increasing
    J = np.array([x["data"] for x in h[1:]]) + lam0 / 2 * np.array([x["tv"] for x in h[:-1]])
    # (ADMM is not a descent method: the last iterate is the minimum of the second half
    # to 1e-5, and the last 20 iterations move J by less than 2e-4)
    late = J[niter // 2:]
    assert J[-1] < 0.5 * J[0] and J[-1] <= late.min() * (1 + 1e-5)
    assert J[-20:].max() - J[-1] < 2e-4 * J[-1]


PINS = [pin_basis_orientation, pin_frame_times, pin_chunk_order, pin_rho_direction, pin_admm_fixed_point]


# ---------------------------------------------------------------- tests (the real oracle)
def test_basis_orientation_full_dft_basis():
    pin_basis_orientation(O_REAL)


def test_frame_times_are_c_integer_division():
    pin_frame_times(O_REAL)


def test_chunk_order_does_not_matter():
    pin_chunk_order(O_REAL)


def test_rho_rule_balances_residuals():
    pin_rho_direction(O_REAL)


def test_admm_converges_to_its_fixed_point():
    pin_admm_fixed_point(O_REAL)


def test_rho_margin():
    assert O_REAL.rho_margin(0.0, 1.0) == np.inf
    assert abs(O_REAL.rho_margin(10.0, 1.0)) < 1e-12          # exactly at the threshold
    assert abs(O_REAL.rho_margin(1.0, 1.0) - np.log(10)) < 1e-12
