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
    (64, 48, 3, "ramp", 32, "cubic"),          # smallest grids: theta half length 32, rho 128
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


@pytest.mark.gpu
def test_bubbles_fixture_backprojection(torch_cuda):
    """The reference's bubbles fixture (tests/data/Rbub, fbub: Hamming-filtered
    Radon data of three discs, N = 256, 64 angles): the device back-projection
    matches the oracle to 1e-5 and the filtered image to the oracle's own
    fitted error (0.217, angular undersampling)."""
    import os
    torch = torch_cuda
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    N, P0 = 256, 64
    R = np.fromfile(os.path.join(golden, "Rbub"), dtype=np.float32).reshape(N, P0).T.copy()
    f = np.fromfile(os.path.join(golden, "fbub"), dtype=np.float32).reshape(N, N)
    lp, tab = make(N, P0, 1, "none", N / 2, "cubic")
    out = torch.zeros((1, N, N), dtype=torch.float32, device="cuda")
    lp.adjp(out, torch.from_numpy(R[None]).cuda(), 0)
    torch.cuda.synchronize()
    x = out.cpu().numpy()[0].astype(np.float64)
    xm = mirror.adj(tab, R[None])[0]
    ref = f[::-1]
    a = float(np.sum(x * ref) / np.sum(x * x))
    err = float(np.linalg.norm(a * x - ref) / np.linalg.norm(ref))
    print("bubbles: rel to oracle %.3e, fitted error to the image %.4f" % (rel(x, xm), err), flush=True)
    assert rel(x, xm) <= TOL
    assert err < 0.25


@pytest.mark.gpu
@pytest.mark.parametrize("interp", ["cubic", "linear"])
def test_box_stride_rule_keeps_results(torch_cuda, interp):
    """The LDS row stride of the gather boxes (tuning key 9, fixed before the
    tables are built) changes where box values sit in LDS, not the arithmetic:
    fwd, adj and both transposes agree bit for bit between the width|1 rule
    (mask 0) and the bank-spread rule on every gather (mask 15) -- fwd and
    adj bit for bit.  A wider stride can split a tile whose box no longer
    fits, which regroups the transposes' scatter sums: those are held to the
    oracle parity bound, 1e-5 relative (measured 1.1e-6 for adj_T, whose
    outputs carry cancellation)."""
    torch = torch_cuda
    N0, P0, ns = 256, 300, 3
    outs = []
    f = torch.from_numpy(phantom_stack(ns, N0, 12, seed=5)).cuda()
    R = torch.from_numpy(np.random.default_rng(8).standard_normal((ns, P0, N0)).astype(np.float32)).cuda()
    for mask in (0, 15):
        lp = lpTransform(N0, P0, ns, "parzen", N0 / 2 - 3, interp)
        lp.precompute(1)
        lp.initcmem(1, 0, schedule={"boxstride": mask})
        assert lp.clphandle[0].get_tuning(9) == mask
        Ro = torch.zeros((ns, P0, N0), device="cuda")
        fo = torch.zeros((ns, N0, N0), device="cuda")
        fT = torch.zeros((ns, N0, N0), device="cuda")
        RT = torch.zeros((ns, P0, N0), device="cuda")
        lp.fwdp(Ro, f, 0)
        lp.adjp(fo, R, 0)
        lp.fwdp_transpose(fT, R, 0)
        lp.adjp_transpose(RT, f, 0)
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy() for t in (Ro, fo, fT, RT)])
    for name, a, b in zip(("fwd", "adj", "fwd_T", "adj_T"), outs[0], outs[1]):
        d = float(np.max(np.abs(a - b)))
        print("%s: max |stride 0 - stride 15| = %.3e" % (name, d), flush=True)
        if name in ("fwd_T", "adj_T"):
            assert rel(a, b) <= TOL
        else:
            assert d == 0.0


@pytest.mark.gpu
def test_polar_gather_slices_per_block_keep_results(torch_cuda):
    """Slices per block of the log-polar -> polar gather (tuning key 13) only
    regroups work: every point is the same 16-tap sum, so fwd is bit-identical
    for 1, 3 and 8 slices per block (a ragged last group included)."""
    torch = torch_cuda
    N0, P0, ns = 256, 300, 7
    lp = lpTransform(N0, P0, ns, "none", N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    h = lp.clphandle[0]
    f = torch.from_numpy(phantom_stack(ns, N0, 12, seed=4)).cuda()
    outs = []
    try:
        for zg in (1, 3, 8):
            h.set_tuning(13, zg)
            R = torch.zeros((ns, P0, N0), device="cuda")
            lp.fwdp(R, f, 0)
            torch.cuda.synchronize()
            outs.append(R.cpu().numpy())
    finally:
        h.set_tuning(13, 4)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.gpu
def test_direct_image_gather_matches_tiled(torch_cuda):
    """The image -> log-polar gather read from global memory (tuning key 14,
    a measurement path) gives the tiled gather's fwd to rounding (1e-6
    relative: the B-spline weights are formed in a different order)."""
    torch = torch_cuda
    N0, P0, ns = 256, 300, 3
    lp = lpTransform(N0, P0, ns, "none", N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    h = lp.clphandle[0]
    f = torch.from_numpy(phantom_stack(ns, N0, 12, seed=6)).cuda()
    outs = []
    try:
        for d in (0, 1):
            h.set_tuning(14, d)
            R = torch.zeros((ns, P0, N0), device="cuda")
            lp.fwdp(R, f, 0)
            torch.cuda.synchronize()
            outs.append(R.cpu().numpy())
    finally:
        h.set_tuning(14, 0)
    e = rel(outs[1], outs[0])
    print("direct vs tiled image gather: %.2e" % e, flush=True)
    assert e <= 1e-6
