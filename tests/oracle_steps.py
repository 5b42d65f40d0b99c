"""The two phases of the reference's per-chunk ADMM step (admm.cu:74-122) and the
rho rule (rectv.cu:224-232), composed from the pinned FP64 oracle operators, for
comparing the fused GPU kernels one phase at a time; plus the comparison of a
GPU history with an oracle trajectory, which flags near-ties of the rho rule.
Test infrastructure (imports the oracle)."""
import numpy as np

from oracle import rectv_oracle as O

NEAR_TIE = 1e-3  # log-distance of r/s from a threshold of the rho rule below which FP32 may branch differently


def chunk_ext(cur, halo_src, iz, nzp, nch):
    """Chunk iz of `cur` with one halo slab each side taken from `halo_src` (the
    previous iterate; rectv.cu:170-171), slab 0 / nzp+1 unused at the volume ends."""
    lo = iz * nzp
    fext = np.zeros((nzp + 2,) + cur.shape[1:], np.float64)
    fext[1:-1] = cur[lo:lo + nzp]
    if iz > 0:
        fext[0] = halo_src[lo - 1]
    if iz < nch - 1:
        fext[-1] = halo_src[lo + nzp]
    return fext, int(iz > 0), int(iz < nch - 1)


def inner_step(ops, A, fcur, g, psi, mu, rho, nzp):
    """fout = fcur - 1/2 div(rho (grad fcur - psi) + mu) - 1/2 R^*(R fcur - g) per chunk
    (admm.cu:85-97), halos from A."""
    nz = fcur.shape[0]
    nch = nz // nzp
    out = np.empty(fcur.shape, np.float64)
    for iz in range(nch):
        fext, fl, fr = chunk_ext(fcur, A, iz, nzp, nch)
        lo = iz * nzp
        h2 = rho * ops.gradient(fext, fl, fr) - rho * psi[iz] + mu[iz]
        h1 = ops.radonapr(fext[1:-1], 1.0) - g[lo:lo + nzp]
        inner = fext[1:-1].copy()
        ops.divergent(inner, h2, -0.5)
        ops.radonapradj(inner, h1, -0.5)
        out[lo:lo + nzp] = inner
    return out


def prox_step(ops, A, fcur, psi, mu, rho, lam0, lam1, nzp):
    """The closing phase (admm.cu:101-117) with fcur as the new iterate and A the
    previous one, the rho rule and div(mu - rho' psi) for the next iteration.
    Returns dict(psi, mu, dchi, r, s, rdiff, fnorm, tv, rho, rho_next)."""
    nz = fcur.shape[0]
    nch = nz // nzp
    psi_n, mu_n = np.empty(psi.shape), np.empty(mu.shape)
    r = s_raw = tv = 0.0
    for iz in range(nch):
        fe, fl, fr = chunk_ext(fcur, A, iz, nzp, nch)
        ae, _, _ = chunk_ext(A, A, iz, nzp, nch)
        h2 = ops.gradient(fe, fl, fr)
        hs = ops.gradient(ae, fl, fr)
        p = O.solve_reg(h2, mu[iz], lam0, rho)
        psi_n[iz] = p
        mu_n[iz] = mu[iz] + rho * h2 - rho * p
        r += float(np.sum((h2 - p) ** 2))
        s_raw += float(np.sum((hs - h2) ** 2))
        tv += float(np.sum(np.sqrt(np.sum(h2 ** 2, axis=-1))))
    s = rho * rho * s_raw
    rho_next = rho * 2 if r > 10 * s else (rho * 0.5 if s > 10 * r else rho)
    dchi = np.empty(fcur.shape)
    for iz in range(nch):
        lo = iz * nzp
        dchi[lo:lo + nzp] = O.div(np.zeros((nzp,) + fcur.shape[1:]), mu_n[iz] - rho_next * psi_n[iz], 1.0, lam1)
    return dict(psi=psi_n, mu=mu_n, dchi=dchi, r=r, s=s, rdiff=float(np.sum((fcur - A) ** 2)),
                fnorm=float(np.sum(fcur.astype(np.float64) ** 2)), tv=tv, rho=rho, rho_next=rho_next)


def compare_trajectory(hist, ref, rtol, say=print):
    """GPU history rows (r, s, rho, rdiff, fnorm, data, tv) against oracle history
    dicts.  The rho sequences must be equal up to the first near-tie of the rho rule
    (margin < NEAR_TIE), where FP32 and FP64 may legitimately branch apart; returns
    the number of iterations compared (len(ref) when no near-tie occurs)."""
    hist = np.asarray(hist)
    upto = len(ref)
    for k, x in enumerate(ref):
        mg = O.rho_margin(x["r"], x["s"])
        if mg < NEAR_TIE:
            say("  NEAR TIE of the rho rule at iteration %d (margin %.2e): trajectory compared up to it" % (k, mg))
            upto = k + 1
            break
    margins = [O.rho_margin(x["r"], x["s"]) for x in ref[:upto]]
    say("  rho rule margins (min %.3g):" % min(margins), ["%.2g" % m for m in margins])
    for k in range(upto):
        x, h = ref[k], hist[k]
        assert h[2] == x["rho"], ("rho", k, h[2], x["rho"])
        for c, key in ((0, "r"), (1, "s"), (3, "rdiff"), (4, "fnorm"), (5, "data"), (6, "tv")):
            if hist.shape[1] <= c:
                continue
            want = x[key]
            assert abs(h[c] - want) <= rtol * abs(want) + 1e-30, (key, k, h[c], want)
    return upto
