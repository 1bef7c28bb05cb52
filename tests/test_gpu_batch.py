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


def test_multiscan_replay_matches_mirror_per_scan(torch_cuda):
    """The uneven multi-scan workload's geometries (N0 in {512, 1024, 2048},
    P0 in {360, 720, 1500}, one slice each) in one captured graph: after
    replay() every scan's sinogram and back-projection equal the oracle
    mirror's (relative L2 <= 1e-5)."""
    torch = torch_cuda
    from oracle import mirror
    from lprec_amd.batch import Scan
    scans = [Scan(n, p, 1) for n in (512, 1024, 2048) for p in (360, 720, 1500)]
    ms = MultiScan(scans)
    imgs = [phantom_stack(1, s.N0, 12, seed=500 + i) for i, s in enumerate(scans)]
    for i, im in enumerate(imgs):
        ms.scan_views(i)[0].copy_(torch.from_numpy(im))
    ms.capture()
    for g in ms.groups:   # the replay must overwrite these
        g.R.fill_(7.0)
        g.back.fill_(-3.0)
    torch.cuda.synchronize()
    ms.replay()
    torch.cuda.synchronize()
    worst = 0.0
    for i, s in enumerate(scans):
        g = ms.groups[ms.where[i][0]]
        tab = mirror.Tables(g.lp.glpars, g.lp.fwdparsi, g.lp.fwdparamsf, g.lp.adjparsi, g.lp.adjparamsf)
        _, Rg, bg = ms.scan_views(i)
        Rm = mirror.fwd(tab, imgs[i])
        bm = mirror.adj(tab, Rg.cpu().numpy())
        ef, ea = rel(Rg.cpu(), Rm), rel(bg.cpu(), bm)
        print("scan N0 %d P0 %d: fwd %.2e adj %.2e" % (s.N0, s.Nproj, ef, ea), flush=True)
        worst = max(worst, ef, ea)
    assert worst <= 1e-5


def test_graphed_em_matches_fp64_oracle_em_at_2048(torch_cuda):
    """EM at the batched EM workload's geometry (N0 = 2048, 2048 angles, 30
    ellipses, 5e3 photons), 2 slices, one iteration per graph replay, against
    oracle.solvers.em on the FP64 mirror after each of 3 iterations
    (relative L2 <= 1e-5 per iterate)."""
    torch = torch_cuda
    from oracle import mirror, solvers
    from lprec_amd import lpmethods
    from lprec_amd.phantom import scan
    N0, P0, ns = 2048, 2048, 2
    lp = lpTransform(N0, P0, ns, "none", N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    _, g = scan(ns, N0, P0, 30, seed=55, photons=5e3)
    g = np.maximum(g, 0).astype(np.float32)
    init = np.full((ns, N0, N0), 1e-3, np.float32)
    tab = mirror.Tables(lp.glpars, lp.fwdparsi, lp.fwdparamsf, lp.adjparsi, lp.adjparamsf)
    _, kept = solvers.em(lambda f: mirror.fwd(tab, f), lambda R: mirror.adj(tab, R), init, g, 3,
                         float(np.float32(1e-3)), keep=(1, 2, 3))
    em = lpmethods.GraphedEM(lp, torch.from_numpy(init).cuda(), torch.from_numpy(g).cuda(), 1e-3, 0)
    em.capture()
    worst = 0.0
    for it in (1, 2, 3):
        em.step()
        torch.cuda.synchronize()
        e = rel(em.recon.cpu().numpy(), kept[it])
        print("EM iteration %d: relative difference to the FP64 oracle %.2e" % (it, e), flush=True)
        worst = max(worst, e)
    assert worst <= 1e-5
