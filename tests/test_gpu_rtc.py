"""Row and column passes compiled at run time (hipRTC) for transform lengths
with odd factors, and the filtered sinogram pair pass at every length,
against the generic kernels they replace (tuning key 10 = 0) and against the
oracle.

The compiled passes run the same Stockham stages with the same butterflies
and twiddle products; only the index arithmetic is resolved at compile time,
so fwd and adj agree with the generic kernels to FP32 rounding (bound 1e-6
relative) and with the FP64 oracle mirror at the parity bound 1e-5.
"""
import numpy as np
import pytest

from lprec_amd import lpTransform
from lprec_amd.phantom import phantom_stack

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


CASES = [
    # N0, P0, slices, filter, cor: padded N, its theta half length h and rho length 2N
    (512, 400, 3, "parzen", 253, "cubic"),     # N = 528: h = 264 = 2^3 3 11, 1056 = 2^5 3 11
    (224, 90, 2, "none", 112, "cubic"),        # N = 224: h = 112 = 2^4 7, 448 = 2^6 7
    (2048, 300, 2, "ramp", 1024.912, "cubic"), # N = 2064: h = 1032 = 2^3 3 43, 4128 = 2^5 3 43
    (250, 300, 2, "hamming", 127.3, "linear"), # N = 256: h = 128 is below the radix-16 row kernels' range
    (1024, 360, 2, "none", 512, "cubic"),      # N = 1024: the radix-16 kernels, no run-time passes
    (1024, 360, 2, "shepp-logan", 512, "cubic"),  # N = 1024 with a filter: the compiled sinogram pair pass only
    (4096, 360, 1, "parzen", 2048, "cubic"),   # N = 4096: rho 8192 (above the radix-16 column kernels), filter
]


def _covered(n):
    """Lengths the radix-16 register kernels handle (16 MID 16, MID <= 16)."""
    return 256 <= n <= 4096 and n & (n - 1) == 0


def _run(lp, f, R, torch):
    ns, P0, N0 = R.shape
    Ro = torch.zeros((ns, P0, N0), device="cuda")
    fo = torch.zeros((ns, N0, N0), device="cuda")
    lp.fwdp(Ro, torch.from_numpy(f).cuda(), 0)
    lp.adjp(fo, torch.from_numpy(R).cuda(), 0)
    torch.cuda.synchronize()
    return Ro.cpu().numpy(), fo.cpu().numpy()


@pytest.mark.parametrize("N0,P0,ns,filt,cor,interp", CASES)
def test_runtime_compiled_passes_match_generic_kernels(torch_cuda, N0, P0, ns, filt, cor, interp):
    torch = torch_cuda
    print("case", N0, P0, ns, filt, cor, interp, flush=True)
    f = phantom_stack(ns, N0, 12, seed=N0 + 1)
    R = np.random.default_rng(N0).standard_normal((ns, P0, N0)).astype(np.float32)
    lp = lpTransform(N0, P0, ns, filt, cor, interp)
    lp.precompute(1)
    lp.initcmem(1, 0)
    h = lp.clphandle[0]
    N = h.info()["N"]
    # the filtered sinogram pair pass is compiled at every length
    expect = 0 if _covered(N // 2) and _covered(2 * N) and filt.lower() == "none" else 1
    assert h.get_tuning(10) == expect, "run-time compiled passes loaded: %d, expected %d" % (h.get_tuning(10), expect)
    a = _run(lp, f, R, torch)
    h.set_tuning(10, 0)
    assert h.get_tuning(10) == 0
    b = _run(lp, f, R, torch)
    h.set_tuning(10, 1)
    ef, ea = rel(a[0], b[0]), rel(a[1], b[1])
    print("N %d: compiled vs generic passes fwd %.2e adj %.2e" % (N, ef, ea), flush=True)
    assert ef <= 1e-6 and ea <= 1e-6


def test_runtime_compiled_passes_match_oracle_at_2064(torch_cuda):
    """FBP (Parzen) and fwd at the centre-offset geometry N0 = 2048 -> N = 2064
    through the compiled passes against the FP64 oracle mirror, 1 slice."""
    torch = torch_cuda
    from oracle import mirror
    N0, P0 = 2048, 1500
    lp = lpTransform(N0, P0, 1, "parzen", 1024.912, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    assert lp.clphandle[0].get_tuning(10) == 1
    f = phantom_stack(1, N0, 20, seed=3)
    R = np.random.default_rng(5).standard_normal((1, P0, N0)).astype(np.float32)
    Rg, fg = _run(lp, f, R, torch)
    tab = mirror.Tables(lp.glpars, lp.fwdparsi, lp.fwdparamsf, lp.adjparsi, lp.adjparamsf)
    ef, ea = rel(Rg, mirror.fwd(tab, f)), rel(fg, mirror.adj(tab, R))
    print("N 2064 compiled passes vs oracle: fwd %.2e adj %.2e" % (ef, ea), flush=True)
    assert ef <= 1e-5 and ea <= 1e-5


def test_compiled_pass_variants_agree(torch_cuda):
    """Every variant of the compiled passes (tuning key 12: prefetch, per-slice
    index arithmetic, waves per SIMD, odd-radix items of 4 or 8 output pairs)
    runs the same butterflies in the same order: fwd and adj agree with the
    default variant to 1e-6 (N = 528)."""
    torch = torch_cuda
    N0, P0, ns = 512, 400, 3
    f = phantom_stack(ns, N0, 12, seed=7)
    R = np.random.default_rng(9).standard_normal((ns, P0, N0)).astype(np.float32)
    lp = lpTransform(N0, P0, ns, "parzen", 253, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    h = lp.clphandle[0]
    default = h.get_tuning(12)
    ref = _run(lp, f, R, torch)
    try:
        for var in (0, 1, 3, 6, 15):
            h.set_tuning(12, var)
            assert h.get_tuning(10) == 1
            a = _run(lp, f, R, torch)
            ef, ea = rel(a[0], ref[0]), rel(a[1], ref[1])
            print("variant %2d vs %d: fwd %.2e adj %.2e" % (var, default, ef, ea), flush=True)
            assert ef <= 1e-6 and ea <= 1e-6
    finally:
        h.set_tuning(12, default)
