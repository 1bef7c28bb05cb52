"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle mirror.

Both sides consume the same wire-format tables (produced by the native
precompute), so the comparison isolates the device pipeline.  Tolerance:
relative L2 <= 1e-5 (FP32 device arithmetic vs FP64 mirror).
"""
import numpy as np
import pytest

from oracle import mirror
from lprec_amd import lpTransform
from lprec_amd.phantom import phantom_stack

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def make(N0, P0, ns, filt, cor, interp):
    lp = lpTransform(N0, P0, ns, filt, cor, interp)
    lp.precompute(1)
    lp.initcmem(1, 0)
    tab = mirror.Tables(lp.glpars, lp.fwdparsi, lp.fwdparamsf, lp.adjparsi, lp.adjparamsf)
    return lp, tab


CASES = [
    # N0, P0, slices, filter, cor, interp
    (256, 384, 4, "none", 128, "cubic"),       # config-1 geometry
    (256, 384, 2, "none", 128, "linear"),
    (250, 300, 2, "ramp", 127.3, "cubic"),     # centre offset -> N = 256, os = 1
    (512, 384, 2, "hamming", 256, "cubic"),    # fixture geometry, os = 2
    (528, 401, 1, "parzen", 261, "cubic"),     # N = 528 = 2^4 3 11 (generic radix), odd P0
    (224, 34, 2, "none", 112, "cubic"),        # os = 10 and 224 = 2^5 7: the foam fixture's regime
]


@pytest.mark.gpu
@pytest.mark.parametrize("N0,P0,ns,filt,cor,interp", CASES)
def test_fwd_adj_match_mirror(torch_cuda, N0, P0, ns, filt, cor, interp):
    torch = torch_cuda
    print("case", N0, P0, ns, filt, cor, interp, flush=True)
    lp, tab = make(N0, P0, ns, filt, cor, interp)
    f = phantom_stack(ns, N0, 12, seed=N0 + P0)
    rng = np.random.default_rng(7)
    R = rng.standard_normal((ns, P0, N0)).astype(np.float32)
    fg = torch.from_numpy(f).cuda()
    Rg = torch.from_numpy(R).cuda()
    Rout = torch.zeros((ns, P0, N0), dtype=torch.float32, device="cuda")
    fout = torch.zeros((ns, N0, N0), dtype=torch.float32, device="cuda")
    lp.fwdp(Rout, fg, 0)
    lp.adjp(fout, Rg, 0)
    torch.cuda.synchronize()
    Rm = mirror.fwd(tab, f)
    fm = mirror.adj(tab, R)
    ef, ea = rel(Rout.cpu().numpy(), Rm), rel(fout.cpu().numpy(), fm)
    print("rel fwd %.3e adj %.3e" % (ef, ea), flush=True)
    assert ef <= TOL and ea <= TOL
