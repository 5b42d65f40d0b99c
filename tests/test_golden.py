"""The oracle reproduces its committed golden vectors (tests/golden/oracle_small.npz,
made by tests/golden/make_golden.py); the GPU path matches the same vectors."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402

FIX = os.path.join(HERE, "golden", "oracle_small.npz")


def load():
    z = np.load(FIX, allow_pickle=False)
    inp = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    out = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    for k in ("n", "nt", "m", "nz", "nzp", "niter", "titer"):
        inp[k] = int(inp[k])
    return inp, out


def test_oracle_matches_golden():
    inp, out = load()
    c = make_golden.case()
    for k in ("f", "g", "fext", "theta", "phi"):
        assert np.array_equal(c[k], inp[k]), k
    got = make_golden.compute(inp)
    for k, v in out.items():
        np.testing.assert_allclose(got[k], v, rtol=1e-11, atol=1e-12 * np.abs(v).max(), err_msg=k)


@pytest.mark.gpu
def test_gpu_matches_golden(gpu):
    from rectv_gpu_amd import Solver
    inp, out = load()
    n, nt, m, nz, nzp = inp["n"], inp["nt"], inp["m"], inp["nz"], inp["nzp"]
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    with Solver(n, nt, m, nz, nzp, 1) as s:
        s.set_inputs(inp["g"], inp["theta"], inp["phi"], float(inp["center"]))
        assert rel(s.apply_radon(inp["f"]), out["radon"]) < 1e-5
        assert rel(s.apply_radon_adj(inp["g"]), out["radon_adj"]) < 1e-5
        h2 = s.apply_gradient(inp["fext"], 0, float(inp["lambda1"]))
        assert rel(h2, out["grad_chunk0"]) < 1e-5
        rec = s.recon(inp["g"], inp["theta"], inp["phi"], rot_center=float(inp["center"]),
                      lambda0=float(inp["lambda0"]) / float(inp["lambda1"]), lambda1=float(inp["lambda1"]),
                      niter=inp["niter"], titer=inp["titer"])
        assert rel(rec, out["run"]) < 1e-4
    with Solver(n, nt, m, 1, 1, 1) as s1:
        res = s1.adjoint_tests(inp["g"][:1], inp["theta"], inp["phi"], float(inp["lambda1"]),
                               rot_center=float(inp["center"]))
    assert np.all(np.abs(res - out["check_adjoints"]) <= 1e-5 * np.abs(out["check_adjoints"]))
