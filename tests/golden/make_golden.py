"""Writes tests/golden/oracle_small.npz: inputs and FP64 oracle outputs of every
operator on a small case (n=32, 24 projections over two half-turns, M=4 Fourier
basis, nz=2, nzp=1, centre n/2+0.5).  The reference ships no golden vectors and
cannot run here (no CUDA), so these vectors pin the oracle against regressions
and give the GPU tests a fixture that needs no oracle run.

  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import rectv_oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_small.npz")


def case():
    n, nt, m, nz, nzp = 32, 24, 4, 2, 1
    theta = (np.arange(nt) * np.pi / 12).astype(np.float32)
    phi = O.takephi(nt, m)
    rng = np.random.default_rng(2024)
    f = rng.standard_normal((nz, m, n, n)).astype(np.float32)
    g = rng.random((nz, nt, n)).astype(np.float32)
    fext = rng.standard_normal((nzp + 2, m, n, n)).astype(np.float32)
    return dict(n=n, nt=nt, m=m, nz=nz, nzp=nzp, theta=theta, phi=phi, f=f, g=g, fext=fext,
                center=np.float32(n / 2 + 0.5), lambda1=np.float32(4.0), lambda0=np.float32(4e-3),
                niter=3, titer=2)


def compute(c):
    n, nt, m, nz, nzp = c["n"], c["nt"], c["m"], c["nz"], c["nzp"]
    ops = O.Operators(n, nt, m, c["theta"], c["phi"], float(c["center"]), float(c["lambda1"]))
    out = {}
    out["radon"] = ops.radonapr(c["f"].astype(np.float64))
    out["radon_adj"] = ops.radonapradj(np.zeros((nz, m, n, n)), c["g"].astype(np.float64))
    out["grad_chunk0"] = O.gradf(O.extendf(c["fext"].astype(np.float64), 0, 1), float(c["lambda1"]))
    out["div_of_grad"] = O.div(np.zeros((nzp, m, n, n)), out["grad_chunk0"], 1.0, float(c["lambda1"]))
    out["check_adjoints"] = O.check_adjoints(c["g"][:1], c["theta"], c["phi"], n, nt, m, 1, 1,
                                             float(c["lambda1"]), float(c["center"]))
    h = []
    out["run"] = O.run(c["g"], c["theta"], c["phi"], n, nt, m, nz, nzp, float(c["center"]), float(c["lambda0"]),
                       float(c["lambda1"]), c["niter"], c["titer"], history=h)
    out["run_rho"] = np.array([x["rho"] for x in h])
    out["run_rdiff"] = np.array([x["rdiff"] for x in h])
    return out


if __name__ == "__main__":
    c = case()
    o = compute(c)
    np.savez_compressed(OUT, **{("in_" + k): np.asarray(v) for k, v in c.items()},
                        **{("out_" + k): v for k, v in o.items()})
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
