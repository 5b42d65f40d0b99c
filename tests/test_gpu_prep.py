"""GPU pre-processing against the FP64 oracle (relative L2): dark/flat normalisation
with -log (1e-6), stripe removal (1e-6), Paganin phase retrieval (1e-5), including
an odd projection height whose padded transform (45 rows -> 63 = 7 * 9) uses
radices 3 and 7."""
import sys

import numpy as np
import pytest

from oracle import rectv_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("nt,nz,n", [(40, 3, 64), (17, 5, 100)])
def test_darkflat_parity(gpu, nt, nz, n):
    from rectv_gpu_amd import prep
    rng = np.random.default_rng(nt + n)
    p = rng.uniform(0.05, 2.5, (nt, nz, n))
    dark = rng.uniform(90, 110, (4, nz, n))
    flat = rng.uniform(900, 1100, (6, nz, n))
    data = (np.exp(-p) * (1e-3 + flat.mean(0) - dark.mean(0)) + dark.mean(0)).astype(np.float32)
    say("case dark/flat", nt, nz, n)
    out = prep.darkflat(data, dark.astype(np.float32), flat.astype(np.float32))
    ref = O.darkflat_log(data, dark.astype(np.float32), flat.astype(np.float32))
    e = rel(out, ref)
    say("  rel L2", e)
    assert out.shape == (nz, nt, n) and e < 1e-6


@pytest.mark.parametrize("nz,nt,n,beta,ms", [(3, 50, 64, 0.022, 1), (2, 33, 256, 0.05, 2), (2, 9, 32, 0.022, 0)])
def test_stripe_removal_parity(gpu, nz, nt, n, beta, ms):
    from rectv_gpu_amd import prep
    rng = np.random.default_rng(n + nt)
    g = rng.uniform(0, 1, (nz, nt, n)).astype(np.float32)
    g[:, :, n // 3] += 2.0
    g[:, :, n // 2 + 5] -= 1.0
    say("case stripe removal", nz, nt, n, beta, ms)
    out = prep.remove_stripe_ti(g, beta, ms)
    ref = O.remove_stripe_ti(g.astype(np.float64), beta, ms)
    e = rel(out, ref)
    say("  rel L2", e)
    assert e < 1e-6


@pytest.mark.parametrize("nt,ny,nz,alpha", [(3, 32, 64, 5e-4), (2, 45, 40, 1e-3), (2, 100, 90, 5e-4)])
def test_paganin_parity(gpu, nt, ny, nz, alpha):
    from rectv_gpu_amd import prep
    px, dist, en = 3.13e-4, 25.0, 25.0
    rng = np.random.default_rng(ny + nz)
    yy, xx = np.mgrid[0:ny, 0:nz]
    base = 0.6 + 0.3 * np.exp(-((yy - ny / 2) ** 2 + (xx - nz / 3) ** 2) / 30.0)
    proj = (base[None] + 0.05 * rng.standard_normal((nt, ny, nz))).astype(np.float32)
    py, pz = O.paganin_pad_width(ny, px, dist, en), O.paganin_pad_width(nz, px, dist, en)
    say("case Paganin", nt, ny, nz, "padded", ny + 2 * py, nz + 2 * pz)
    out = prep.retrieve_phase(proj, px, dist, en, alpha)
    ref = O.retrieve_phase(proj.astype(np.float64), px, dist, en, alpha)
    e = rel(out, ref)
    say("  rel L2", e)
    assert e < 1e-5
