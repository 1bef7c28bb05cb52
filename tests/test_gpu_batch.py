"""Uneven multi-scan batches in one graph launch, the autotuner, and FBP at
the full centre-offset geometry (GPU).

* every scan's fwd/adj result from the captured graph equals a stand-alone
  lpTransform call on that scan alone (relative 1e-6: the same kernels on
  the same slices), on three replays with the program's own eager work and
  input changes between them;
* autotuned schedules give the same results as the default schedule;
* FBP (ramp filter) at N0 = 2048, 1500 angles, cor = 1021.6 (padded N = 2064,
  FFT lengths 2064 = 2^4 3 43 and 4128) against the oracle mirror, 1 slice.
"""
import numpy as np
import pytest

from lprec_amd import lpTransform
from lprec_amd.batch import MultiScan, random_scans
from lprec_amd.phantom import phantom_stack

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_multiscan_graph_matches_single_scan_calls(torch_cuda):
    torch = torch_cuda
    scans = random_scans(12, seed=3, sizes=(128, 256), angles=(90, 180, 301), slices=(1, 4))
    print("scans", [(s.N0, s.Nproj, s.nslices) for s in scans], flush=True)
    ms = MultiScan(scans)
    assert len(ms.groups) <= 6
    rng = np.random.default_rng(0)
    imgs = [phantom_stack(s.nslices, s.N0, 6, seed=100 + i) for i, s in enumerate(scans)]
    for i, im in enumerate(imgs):
        ms.scan_views(i)[0].copy_(torch.from_numpy(im))
    ms.capture()
    singles = []
    for i, s in enumerate(scans):
        lp = lpTransform(s.N0, s.Nproj, s.nslices, "none", s.cor, "cubic")
        lp.precompute(1)
        lp.initcmem(1, 0)
        singles.append(lp)
    for rep in range(3):
        if rep == 2:   # new inputs between replays
            for i, s in enumerate(scans):
                imgs[i] = (imgs[i] * (1 + 0.5 * rng.random())).astype(np.float32)
                ms.scan_views(i)[0].copy_(torch.from_numpy(imgs[i]))
        for g in ms.groups:   # eager work on the outputs between replays
            g.R.fill_(7.0)
            g.back.zero_()
        torch.cuda.synchronize()
        ms.replay()
        torch.cuda.synchronize()
        worst = 0.0
        for i, s in enumerate(scans):
            f = torch.from_numpy(imgs[i]).cuda()
            R = torch.zeros((s.nslices, s.Nproj, s.N0), device="cuda")
            back = torch.zeros((s.nslices, s.N0, s.N0), device="cuda")
            singles[i].fwdp(R, f, 0)
            singles[i].adjp(back, R, 0)
            _, Rg, bg = ms.scan_views(i)
            worst = max(worst, rel(Rg.cpu(), R.cpu()), rel(bg.cpu(), back.cpu()))
        print("replay %d: worst relative difference %.2e" % (rep, worst), flush=True)
        assert worst <= 1e-6


def test_autotuned_schedule_keeps_results(torch_cuda):
    torch = torch_cuda
    from lprec_amd.tune import autotune
    N0, P0, ns = 512, 400, 6
    lp = lpTransform(N0, P0, ns, "none", N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    f = torch.from_numpy(phantom_stack(ns, N0, 8, seed=9)).cuda()
    R0 = torch.zeros((ns, P0, N0), device="cuda")
    b0 = torch.zeros((ns, N0, N0), device="cuda")
    lp.fwdp(R0, f, 0)
    lp.adjp(b0, R0, 0)
    best, ms, tried = autotune(lp, 0, reps=2, sweeps=1)
    print("autotune: %d schedules, best %s at %.3f ms" % (len(tried), best, ms), flush=True)
    assert len(tried) > 5
    R1 = torch.zeros_like(R0)
    b1 = torch.zeros_like(b0)
    lp.fwdp(R1, f, 0)
    lp.adjp(b1, R1, 0)
    assert rel(R1.cpu(), R0.cpu()) <= 1e-6 and rel(b1.cpu(), b0.cpu()) <= 1e-6


def test_fbp_full_size_centre_offset_matches_mirror(torch_cuda):
    torch = torch_cuda
    from oracle import mirror
    from lprec_amd.phantom import scan
    N0, P0, cor = 2048, 1500, 1021.6
    lp = lpTransform(N0, P0, 1, "ramp", cor, "cubic")
    lp.precompute(0)
    lp.initcmem(0, 0)
    assert lp.clphandle[0].info()["N"] == 2064
    _, R = scan(1, N0, P0, 20, seed=2064, photons=1e4, cor=cor)
    rec = torch.zeros((1, N0, N0), device="cuda")
    lp.adjp(rec, torch.from_numpy(R).cuda(), 0)
    tab = mirror.Tables(lp.glpars, None, None, lp.adjparsi, lp.adjparamsf)
    e = rel(rec.cpu().numpy(), mirror.adj(tab, R))
    print("FBP N=2064 vs mirror: %.2e" % e, flush=True)
    assert e <= 1e-5


def test_graphed_em_matches_eager_every_replay(torch_cuda):
    """One EM iteration captured in a HIP graph and replayed 6 times, with
    eager work and a synchronisation between replays, against the same
    iterations run eagerly (and against lpmethods.em)."""
    torch = torch_cuda
    from lprec_amd import lpTransform, lpmethods
    from lprec_amd.phantom import scan
    N0, P0, ns = 256, 384, 4
    lp = lpTransform(N0, P0, ns, "none", N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    _, g = scan(ns, N0, P0, 12, seed=5, photons=5e3)
    tomo = torch.from_numpy(np.maximum(g, 0)).cuda()
    init = torch.full((ns, N0, N0), 1e-3, device="cuda")
    graphed = lpmethods.GraphedEM(lp, init, tomo, 1e-3, 0)
    eager = lpmethods.GraphedEM(lp, init, tomo, 1e-3, 0)
    graphed.capture()
    worst = 0.0
    for it in range(6):
        graphed.step()
        torch.cuda.synchronize()
        eager.step()
        torch.cuda.synchronize()
        a, b = graphed.recon.double(), eager.recon.double()
        err = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        worst = max(worst, err)
        print("replay %d: relative difference to eager %.2e" % (it + 1, err), flush=True)
    ref = lpmethods.em(lp, init, tomo, 6, 1e-3, 0)
    e_ref = float(torch.linalg.norm(graphed.recon.double() - ref.double()) / torch.linalg.norm(ref.double()))
    print("graphed vs lpmethods.em after 6 iterations: %.2e" % e_ref, flush=True)
    assert worst < 1e-5
    assert e_ref < 1e-5
