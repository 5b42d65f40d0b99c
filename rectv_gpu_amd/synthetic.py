"""Synthetic moving-sphere phantoms and their analytic projections.

Geometry convention of the solver: detector coordinate s_c = s - n/2 of the ray
x_c sin(theta) + y_c cos(theta) = s_c (x_c = x - n/2 column, y_c = y - n/2 row),
rotation centre at n/2.  Data are scaled by sqrt(m)/sqrt(n*ntheta), which makes
a static object of density d reconstruct to d in every frame under the Fourier
basis.
"""
import numpy as np


def spheres(n, nz, count, seed, rmin=4.0, rmax=12.0, vmax=0.5, nrot=1, oscillate=False):
    """Random spheres: centre inside the inscribed circle, z in [0, nz), radius U[rmin, rmax],
    density U[0.2, 1], drift U[-vmax, vmax] px per rotation, optional density oscillation
    with period U[1, nrot] rotations."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        r = rng.uniform(rmin, rmax)
        while True:
            cx, cy = rng.uniform(-n / 2 + r + 2, n / 2 - r - 2, 2)
            if cx * cx + cy * cy < (n / 2 - r - 2) ** 2:
                break
        cz = rng.uniform(0, nz)
        d = rng.uniform(0.2, 1.0)
        vx, vy = rng.uniform(-vmax, vmax, 2)
        per = rng.uniform(1, max(nrot, 1)) if oscillate else 0.0
        amp = rng.uniform(0.2, 0.5) if oscillate else 0.0
        out.append((cx + n / 2, cy + n / 2, cz, r, d, vx, vy, per, amp))
    return out


def project(sph, n, nz, theta, ntheta_rot, m, noise=0.0, seed=0):
    """Analytic projections [nz, ntheta, n] float32 of moving spheres at angles theta
    (one time step per projection; drift in px per ntheta_rot projections)."""
    nt = len(theta)
    th = np.asarray(theta, np.float64)
    sn, cs = np.sin(th), np.cos(th)
    k = np.arange(nt, dtype=np.float64)
    data = np.zeros((nz, nt, n), np.float64)
    rows = np.arange(nt)[:, None]
    for (cx, cy, cz, r, d, vx, vy, per, amp) in sph:
        t = k / ntheta_rot
        px = cx + vx * t - n / 2
        py = cy + vy * t - n / 2
        p = px * sn + py * cs                              # projected centre, detector coords
        dens = d * (1 + amp * np.sin(2 * np.pi * t / per)) if per > 0 else np.full(nt, d)
        W = int(2 * np.ceil(r) + 3)
        s0 = np.floor(p + n / 2 - r).astype(np.int64) - 1
        sidx = s0[:, None] + np.arange(W)[None, :]         # [nt, W]
        valid = (sidx >= 0) & (sidx < n)
        sc = sidx - n / 2 - p[:, None]
        z0, z1 = max(0, int(np.floor(cz - r))), min(nz - 1, int(np.ceil(cz + r)))
        sidc = np.clip(sidx, 0, n - 1)
        for z in range(z0, z1 + 1):
            rz2 = r * r - (z - cz) ** 2
            if rz2 <= 0:
                continue
            ch = 2 * np.sqrt(np.maximum(rz2 - sc * sc, 0.0)) * dens[:, None]
            ch[~valid] = 0
            data[z][rows, sidc] += ch
    data *= np.sqrt(m) / np.sqrt(n * nt)
    if noise > 0:
        rng = np.random.default_rng(seed)
        rms = np.sqrt(np.mean(data ** 2))
        data += rng.normal(0, noise * rms, data.shape)
    return data.astype(np.float32)


def workload(cfg, seed=0):
    """Inputs of the named configurations: returns dict(n, nz, nzp, ntheta, m, theta, phi, data, ...)."""
    from .basis import angles, takephi, realbasis
    if cfg == "small":      # N=128, Nz=8, Nrot=4 x 32, M=4 real basis, 20 spheres, 1% noise
        n, nz, nzp, nrot, ntr, m, cnt, nz_noise, niter = 128, 8, 8, 4, 32, 4, 20, 0.01, 32
        phi = realbasis(nrot * ntr, m)
    elif cfg == "mid":      # N=256, Nz=32, Nzp=8, Nrot=4 x 100, M=8, 60 spheres, 2% noise
        n, nz, nzp, nrot, ntr, m, cnt, nz_noise, niter = 256, 32, 8, 4, 100, 8, 60, 0.02, 64
        phi = takephi(nrot * ntr, m)
    elif cfg == "full":     # N=1024, Nz=64, Nzp=16, Nrot=8 x 200, M=16, 400 spheres, 2% noise
        n, nz, nzp, nrot, ntr, m, cnt, nz_noise, niter = 1024, 64, 16, 8, 200, 16, 400, 0.02, 128
        phi = takephi(nrot * ntr, m)
    else:
        raise ValueError(cfg)
    theta = angles(ntr, nrot)
    sph = spheres(n, nz, cnt, seed, nrot=nrot, oscillate=(cfg == "full"))
    data = project(sph, n, nz, theta, ntr, m, noise=nz_noise, seed=seed)
    return dict(n=n, nz=nz, nzp=nzp, ntheta=nrot * ntr, m=m, theta=theta, phi=phi, data=data,
                niter=niter, lambda0=1e-3, lambda1=4.0, center=n / 2, nrot=nrot, ntheta_rot=ntr)


def batch_specs(count=64, seed0=1, rng_seed=4):
    """Sizes of the batched-instances mix: N from {128, 256, 384, 512}, Nz from
    U{4..32}, projections per rotation from {64, 100, 128}, rotations from
    {2, 4, 8}, M from {4, 8, 16}; job i uses phantom seed seed0 + i; Nzp = Nz."""
    rng = np.random.default_rng(rng_seed)
    out = []
    for i in range(count):
        nz = int(rng.integers(4, 33))
        out.append(dict(n=int(rng.choice([128, 256, 384, 512])), nz=nz, nzp=nz,
                        ntr=int(rng.choice([64, 100, 128])), nrot=int(rng.choice([2, 4, 8])),
                        m=int(rng.choice([4, 8, 16])), seed=seed0 + i))
    return out


def serve_specs(count=64, seed0=1001, rng_seed=5, big_frac=0.05):
    """Sizes of the serving mix: the batched mix plus a fraction big_frac of
    N=1024, Nz=32, M=16 jobs (4 rotations x 200 projections)."""
    rng = np.random.default_rng(rng_seed)
    base = batch_specs(count, seed0, rng_seed + 100)
    for i, sp in enumerate(base):
        if rng.random() < big_frac:
            base[i] = dict(n=1024, nz=32, nzp=32, ntr=200, nrot=4, m=16, seed=sp["seed"])
    return base


def make_job(spec, niter, titer, spheres_count=30, noise=0.02, tol=None):
    """A batch.Job of the given sizes: moving-sphere phantom, Fourier basis,
    lambda0=1e-3, lambda1=4, centre n/2."""
    from .basis import angles, takephi
    from .batch import Job
    n, nz, ntr, nrot, m = spec["n"], spec["nz"], spec["ntr"], spec["nrot"], spec["m"]
    theta = angles(ntr, nrot)
    sph = spheres(n, nz, spheres_count, spec["seed"], nrot=nrot)
    data = project(sph, n, nz, theta, ntr, m, noise=noise, seed=spec["seed"])
    return Job(data, theta, takephi(nrot * ntr, m), lambda0=1e-3, lambda1=4.0, niter=niter, titer=titer,
               nzp=spec.get("nzp", nz), tol=tol, name="n%d_nz%d_nt%d_m%d_s%d" % (n, nz, nrot * ntr, m, spec["seed"]))
