"""Writes tests/golden/config2_trajectory.npz: the FP64 oracle's trajectory of the
mid-size reconstruction (N=256, Nz=32, Nzp=8, Nrot=4 x 100 = 400 projections,
M=8 Fourier basis, 60 moving spheres, 2% noise, lambda0=1e-3, lambda1=4,
niter=64, titer=4; inputs from rectv_gpu_amd.synthetic.workload("mid")).

Stored: per outer iteration r, s, rho used, rho next, ||f(k+1)-f(k)||^2,
||f(k+1)||^2, the objective terms (1/2||R f - g||^2 of the entering iterate, sum of
gradient norms of the leaving one) and the rho-rule margin; the final
reconstruction as a strided subsample f[::4, :, ::8, ::8] with its full squared
norm and per-(slice, frame) squared norms; checksums of the generated data so
that the test can confirm it regenerated the same inputs.  The run takes about
half an hour on 8 host threads.

  python tests/golden/make_config2.py [threads]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import rectv_oracle as O  # noqa: E402
from rectv_gpu_amd import synthetic  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config2_trajectory.npz")
TITER = 4


def data_checksums(data):
    d = data.astype(np.float64)
    return np.array([d.sum(), (d * d).sum(), d[0, 0].sum(), d[-1, -1].sum()])


def main(threads=4):
    w = synthetic.workload("mid")
    O.set_workers(max(1, (os.cpu_count() or 1) // threads))
    h = []
    t0 = time.time()

    class Log:
        def write(self, s):
            if s.strip():
                sys.stdout.write("%7.1f s  %s\n" % (time.time() - t0, s.strip()))
                sys.stdout.flush()

        def flush(self):
            pass

    f = O.run(w["data"], w["theta"], w["phi"], w["n"], w["ntheta"], w["m"], w["nz"], w["nzp"], w["center"],
              w["lambda0"] * w["lambda1"], w["lambda1"], w["niter"], TITER, history=h, log=Log(), threads=threads)
    col = lambda k: np.array([x[k] for x in h])
    np.savez_compressed(
        OUT, n=w["n"], nz=w["nz"], nzp=w["nzp"], ntheta=w["ntheta"], m=w["m"], niter=w["niter"], titer=TITER,
        lambda0=w["lambda0"], lambda1=w["lambda1"], data_checksums=data_checksums(w["data"]),
        r=col("r"), s=col("s"), rho=col("rho"), rho_next=col("rho_next"), rdiff=col("rdiff"), fnorm=col("fnorm"),
        data_term=col("data"), tv=col("tv"), margin=np.array([O.rho_margin(x["r"], x["s"]) for x in h]),
        f_sub=f[::4, :, ::8, ::8].astype(np.float32), f_norm2=float(np.sum(f * f)),
        f_frame_norm2=np.sum(f * f, axis=(2, 3)))
    print("wrote", OUT, os.path.getsize(OUT), "bytes in %.0f s" % (time.time() - t0))


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 4)
