"""The reference's operator tests run through the HIP path at the reference's
own thresholds (tests/run_test.py:14-20 of the reference).

* test_adj_ptr (tests/test_adj_ptr.py:7-44): N = 512, 768 angles, one slice
  of uniform random data; err0 = |<R f, g> - <f, R* g>| / |<f, R* g>| < 1e-2
  (the reference's fwd/adj pair is not an exact transpose).
* test_adj_many (tests/test_adj_many.py:9-65): the fixtures f (512 x 512) and
  R (384 x 512), 8 slices with slice k scaled by 8 - k; err0 < 5e-3 (pair
  mismatch), err1 = ||R* R_fixture - f|| / ||f|| < 7e-3 (back-projection of
  the fixture sinogram reproduces the fixture image), err2 = ||R R* R_fixture -
  R f|| / ||R f|| < 7e-4.  The device outputs are also compared with the
  oracle mirror slice by slice (relative L2 <= 1e-5).
"""
import os

import numpy as np
import pytest

from oracle import mirror
from lprec_amd import lpTransform

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _dot(a, b):
    return float(np.sum(np.float64(a).ravel() * np.float64(b).ravel()))


def test_adj_ptr(torch_cuda):
    torch = torch_cuda
    N = 512
    Nproj = int(3 * N / 2)
    rng = np.random.default_rng(512)
    f = np.float32(rng.random((1, N, N)))
    R = np.float32(rng.random((1, Nproj, N)))
    lp = lpTransform(N, Nproj, 1, "None", N / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    Rf = torch.zeros((1, Nproj, N), device="cuda")
    fR = torch.zeros((1, N, N), device="cuda")
    lp.fwdp(Rf, torch.from_numpy(f).cuda(), 0)
    lp.adjp(fR, torch.from_numpy(R).cuda(), 0)
    Rf, fR = Rf.cpu().numpy(), fR.cpu().numpy()
    sum1, sum2 = _dot(Rf, R), _dot(fR, f)
    err0 = abs(sum1 - sum2) / abs(sum2)
    print("test_adj_ptr: err0 %.3e (reference bound 1e-2)" % err0, flush=True)
    assert err0 < 1e-2


def test_adj_many(torch_cuda):
    torch = torch_cuda
    N, ns = 512, 8
    Nproj = int(3 * N / 4)
    f = np.fromfile(os.path.join(GOLDEN, "f"), dtype=np.float32).reshape(1, N, N)
    R = np.fromfile(os.path.join(GOLDEN, "R"), dtype=np.float32).reshape(1, Nproj, N)
    scale = (ns - np.arange(ns, dtype=np.float32)).reshape(ns, 1, 1)
    fa = np.float32(f * scale)
    Ra = np.float32(R * scale)
    lp = lpTransform(N, Nproj, ns, "None", N / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    Rf = torch.zeros((ns, Nproj, N), device="cuda")
    fR = torch.zeros((ns, N, N), device="cuda")
    RfR = torch.zeros((ns, Nproj, N), device="cuda")
    lp.fwdp(Rf, torch.from_numpy(fa).cuda(), 0)
    lp.adjp(fR, torch.from_numpy(Ra).cuda(), 0)
    lp.fwdp(RfR, fR, 0)
    Rf, fR, RfR = (np.float64(t.cpu().numpy()) for t in (Rf, fR, RfR))
    sum1, sum2 = _dot(Rf, Ra), _dot(fR, fa)
    err0 = abs(sum1 - sum2) / abs(sum2)
    err1 = np.linalg.norm(fR - fa) / np.linalg.norm(fa)
    err2 = np.linalg.norm(RfR - Rf) / np.linalg.norm(Rf)
    print("test_adj_many: err0 %.3e err1 %.3e err2 %.3e (bounds 5e-3, 7e-3, 7e-4)" % (err0, err1, err2),
          flush=True)
    tab = mirror.Tables(lp.glpars, lp.fwdparsi, lp.fwdparamsf, lp.adjparsi, lp.adjparamsf)
    Rm, fm = mirror.fwd(tab, fa), mirror.adj(tab, Ra)
    em = max(max(np.linalg.norm(Rf[z] - Rm[z]) / np.linalg.norm(Rm[z]),
                  np.linalg.norm(fR[z] - fm[z]) / np.linalg.norm(fm[z])) for z in range(ns))
    print("test_adj_many: worst slice vs the oracle mirror %.2e" % em, flush=True)
    assert err0 < 5e-3 and err1 < 7e-3 and err2 < 7e-4
    assert em <= 1e-5
