"""GPU exact transposes (lp_fwd_T / lp_adj_T through the C ABI).

* parity: the HIP transposes against the oracle's FP64 transposes
  (oracle/transpose.py, itself pinned by FP64 dot tests in test_transpose.py)
  on the same tables, relative L2 <= 1e-5;
* the adjoint test on the GPU alone, |<F f, R> - <f, F^T R>| / (||F f|| ||R||)
  <= 1e-6 (and the same for adj), at N = 1024.
"""
import numpy as np
import pytest

from oracle import mirror
from oracle import transpose as tr
from lprec_amd import lpTransform

pytestmark = pytest.mark.gpu

TOL = 1e-5

CASES = [
    (256, 384, 2, "none", 128, "cubic"),
    (256, 384, 2, "none", 128, "linear"),
    (250, 300, 2, "ramp", 127.3, "cubic"),
    (528, 401, 1, "parzen", 261, "cubic"),
    (224, 34, 2, "shepp-logan", 112, "cubic"),
]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("N0,P0,ns,filt,cor,interp", CASES)
def test_transposes_match_oracle(torch_cuda, N0, P0, ns, filt, cor, interp):
    torch = torch_cuda
    print("case", N0, P0, ns, filt, cor, interp, flush=True)
    lp = lpTransform(N0, P0, ns, filt, cor, interp)
    lp.precompute(1)
    lp.initcmem(1, 0)
    tab = mirror.Tables(lp.glpars, lp.fwdparsi, lp.fwdparamsf, lp.adjparsi, lp.adjparamsf)
    rng = np.random.default_rng(N0 + P0)
    R = rng.standard_normal((ns, P0, N0)).astype(np.float32)
    f = rng.standard_normal((ns, N0, N0)).astype(np.float32)
    fT = torch.zeros((ns, N0, N0), device="cuda")
    RT = torch.zeros((ns, P0, N0), device="cuda")
    lp.fwdp_transpose(fT, torch.from_numpy(R).cuda(), 0)
    lp.adjp_transpose(RT, torch.from_numpy(f).cuda(), 0)
    torch.cuda.synchronize()
    e1 = rel(fT.cpu().numpy(), tr.fwd_T(tab, R))
    e2 = rel(RT.cpu().numpy(), tr.adj_T(tab, f, P0))
    print("rel fwd_T %.3e adj_T %.3e" % (e1, e2), flush=True)
    assert e1 <= TOL and e2 <= TOL


@pytest.mark.parametrize("filt", ["none", "parzen"])
def test_adjoint_dot_test(torch_cuda, filt):
    torch = torch_cuda
    N0, P0, ns = 1024, 1024, 2
    lp = lpTransform(N0, P0, ns, filt, N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    g = torch.Generator(device="cuda").manual_seed(5)
    f = torch.rand((ns, N0, N0), device="cuda", generator=g)
    R = torch.rand((ns, P0, N0), device="cuda", generator=g)
    Ff = torch.zeros_like(R)
    FtR = torch.zeros_like(f)
    Af = torch.zeros_like(R)
    AR = torch.zeros_like(f)
    lp.fwdp(Ff, f, 0)
    lp.fwdp_transpose(FtR, R, 0)
    lp.adjp(AR, R, 0)
    lp.adjp_transpose(Af, f, 0)
    d = lambda a, b: float(torch.sum(a.double() * b.double()))
    n = lambda a: float(torch.linalg.norm(a.double()))
    err_f = abs(d(Ff, R) - d(f, FtR)) / (n(Ff) * n(R))
    err_a = abs(d(AR, f) - d(R, Af)) / (n(AR) * n(f))
    # for comparison: the reference pair (adj as the adjoint of fwd)
    err_pair = abs(d(Ff, R) - d(f, AR)) / (n(Ff) * n(R))
    print("dot test fwd %.2e adj %.2e  (reference fwd/adj pair %.2e)" % (err_f, err_a, err_pair), flush=True)
    assert err_f <= 1e-6 and err_a <= 1e-6
