"""Writes tests/golden/mix_samples.npz: FP64 oracle results for a 10% sample of the
batched mix (64 jobs of rectv_gpu_amd.synthetic.batch_specs, niter=64, titer=4)
and of the serving mix (64 jobs of synthetic.serve_specs, niter=64, titer=4,
stopping at the first outer iteration whose relative change is below 1e-4).
The sample is the 7 jobs of each mix with the least oracle work (the oracle at
the larger sizes takes hours).  Per job: the stop iteration, the trajectory
(r, s, rho, ||df||^2, ||f||^2, objective terms, rho-rule margin), the result as a
strided subsample f[::zs, :, ::xs, ::xs] with its squared norm and per-(slice,
frame) squared norms, and checksums of the generated data.

  python tests/golden/make_mix_samples.py [processes]
"""
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mix_samples.npz")
NITER, TITER, TOL, COUNT, SAMPLE = 64, 4, 1e-4, 64, 7


def cost(sp):
    return sp["nz"] * sp["m"] * sp["n"] * (sp["ntr"] * sp["nrot"] * 81 + 40 * sp["n"])


def strides(nz, n):
    return max(1, -(-nz // 4)), max(1, n // 16)


def sample_ids(specs):
    return sorted(sorted(range(len(specs)), key=lambda i: cost(specs[i]))[:SAMPLE])


def data_checksums(data):
    d = data.astype(np.float64)
    return np.array([d.sum(), (d * d).sum(), d[0, 0].sum(), d[-1, -1].sum()])


def solve(args):
    mix, i, sp = args
    from oracle import rectv_oracle as O
    from rectv_gpu_amd import synthetic
    O.set_workers(1)
    j = synthetic.make_job(sp, NITER, TITER)
    t0 = time.time()
    h = []
    f = O.run(j.data, j.theta, j.phi, j.n, j.ntheta, j.m, j.nz, j.nzp, j.rot_center, j.lambda0 * j.lambda1,
              j.lambda1, NITER, TITER, history=h, tol=TOL if mix == "serve" else None)
    zs, xs = strides(j.nz, j.n)
    col = lambda k: np.array([x[k] for x in h])
    out = dict(iterations=len(h), r=col("r"), s=col("s"), rho=col("rho"), rdiff=col("rdiff"), fnorm=col("fnorm"),
               data_term=col("data"), tv=col("tv"), margin=np.array([O.rho_margin(x["r"], x["s"]) for x in h]),
               f_sub=f[::zs, :, ::xs, ::xs].astype(np.float32), f_norm2=float(np.sum(f * f)),
               f_frame_norm2=np.sum(f * f, axis=(2, 3)), data_checksums=data_checksums(j.data))
    print("%s job %d %s: %d iterations in %.0f s" % (mix, i, j.name, len(h), time.time() - t0), flush=True)
    return mix, i, out


def main(procs=8):
    from rectv_gpu_amd import synthetic
    work = []
    for mix, specs in (("batch", synthetic.batch_specs(COUNT)), ("serve", synthetic.serve_specs(COUNT))):
        work += [(mix, i, specs[i]) for i in sample_ids(specs)]
    work.sort(key=lambda w: -cost(w[2]))
    res = {}
    with Pool(procs) as p:
        for mix, i, out in p.imap_unordered(solve, work):
            for k, v in out.items():
                res["%s_%d_%s" % (mix, i, k)] = np.asarray(v)
    np.savez_compressed(OUT, **res)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
