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


@pytest.mark.parametrize("N0,P0,ns,interp", [(1024, 1024, 4, "cubic"), (256, 384, 2, "linear"), (2048, 1500, 2, "cubic")])
def test_separable_polar_transpose_matches_tiled_scatter(torch_cuda, N0, P0, ns, interp):
    """lp_fwd_T with the transpose of the polar gather as two gathers (tuning
    key 7 = 1, the default) against the tiled LDS/global-atomic scatter (key 7
    = 0): the same operator (relative L2 <= 1e-6; the summation orders
    differ)."""
    torch = torch_cuda
    lp = lpTransform(N0, P0, ns, "none", N0 / 2, interp)
    lp.precompute(1)
    lp.initcmem(1, 0)
    h = lp.clphandle[0]
    assert h.get_tuning(7) == 1
    R = torch.from_numpy(np.random.default_rng(N0 + 1).standard_normal((ns, P0, N0)).astype(np.float32)).cuda()
    a = torch.zeros((ns, N0, N0), device="cuda")
    b = torch.full((ns, N0, N0), 5.0, device="cuda")
    try:
        lp.fwdp_transpose(a, R, 0)
        h.set_tuning(7, 0)
        lp.fwdp_transpose(b, R, 0)
    finally:
        h.set_tuning(7, 1)
    torch.cuda.synchronize()
    e = rel(a.cpu().numpy(), b.cpu().numpy().astype(np.float64))
    print("N0 %d %s: separable vs scattered polar transpose %.2e" % (N0, interp, e), flush=True)
    assert e <= 1e-6


@pytest.mark.parametrize("N0,P0,ns,filt,interp", [(1024, 1024, 4, "none", "cubic"), (250, 300, 2, "ramp", "cubic"),
                                                  (256, 384, 2, "none", "linear")])
def test_staged_scatter_flush_matches_atomic_flush(torch_cuda, N0, P0, ns, filt, interp):
    """lp_fwd_T and lp_adj_T with the tiled scatters' boxes staged and summed
    per destination tile (tuning key 8 = 1, the default; the polar transpose
    forced onto the tiled scatter too) against the global-atomic flush (key 8
    = 0): relative L2 <= 1e-6 (the summation orders differ)."""
    torch = torch_cuda
    lp = lpTransform(N0, P0, ns, filt, N0 / 2 if filt == "none" else 127.3, interp)
    lp.precompute(1)
    lp.initcmem(1, 0)
    h = lp.clphandle[0]
    rng = np.random.default_rng(N0 + 2)
    R = torch.from_numpy(rng.standard_normal((ns, P0, N0)).astype(np.float32)).cuda()
    f = torch.from_numpy(rng.standard_normal((ns, N0, N0)).astype(np.float32)).cuda()
    out = {}
    try:
        h.set_tuning(7, 0)
        for staged in (1, 0):
            h.set_tuning(8, staged)
            fT = torch.full((ns, N0, N0), 9.0, device="cuda")
            RT = torch.full((ns, P0, N0), 9.0, device="cuda")
            lp.fwdp_transpose(fT, R, 0)
            lp.adjp_transpose(RT, f, 0)
            torch.cuda.synchronize()
            out[staged] = (fT.cpu().numpy().astype(np.float64), RT.cpu().numpy().astype(np.float64))
    finally:
        h.set_tuning(7, 1)
        h.set_tuning(8, 1)
    e1, e2 = rel(out[1][0], out[0][0]), rel(out[1][1], out[0][1])
    print("N0 %d %s %s: staged vs atomic flush fwd_T %.2e adj_T %.2e" % (N0, filt, interp, e1, e2), flush=True)
    assert e1 <= 1e-6 and e2 <= 1e-6


def test_fwd_transpose_is_bit_reproducible(torch_cuda):
    """The transpose chain of lp_fwd (autograd's backward of radon) sums every
    output in a fixed order -- separable polar transpose, colour-class box
    accumulation, staged flush -- so repeated calls agree bit for bit."""
    torch = torch_cuda
    N0, P0, ns = 1024, 1024, 4
    lp = lpTransform(N0, P0, ns, "none", N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    R = torch.from_numpy(np.random.default_rng(11).standard_normal((ns, P0, N0)).astype(np.float32)).cuda()
    f = torch.from_numpy(np.random.default_rng(12).standard_normal((ns, N0, N0)).astype(np.float32)).cuda()
    outs, outs_a = [], []
    for _ in range(3):
        a = torch.zeros((ns, N0, N0), device="cuda")
        b = torch.zeros((ns, P0, N0), device="cuda")
        lp.fwdp_transpose(a, R, 0)
        lp.adjp_transpose(b, f, 0)
        torch.cuda.synchronize()
        outs.append(a.cpu().numpy())
        outs_a.append(b.cpu().numpy())
    same = all(np.array_equal(outs[0], o) for o in outs[1:])
    da = max(float(np.max(np.abs(outs_a[0] - o))) for o in outs_a[1:])
    print("fwd_T repeated calls bit-identical: %s; adj_T max abs difference %.2e" % (same, da), flush=True)
    assert same
