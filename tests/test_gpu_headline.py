"""GPU parity at the sizes of the measured workloads (through the C ABI).

The headline workload is N0 = 2048, 1500 angles, cor = 1024, cubic, a batch
of 32 slices with the default schedule: the register FFT kernels (forward
rows of h = 1024 complex points in radix 16, inverse rows as 32 x 32 with 32
lanes per row, columns of Nrho = 4096) and the slice grouping of the gathers
and FFT passes.  Checks:

* stage by stage on the 32-slice call (lp_debug_stop): each stage's device
  output against the FP64 restatement of that stage applied to the device's
  own input of the stage, so every kernel is isolated --
  prefilter (stage 2) vs ``mirror.prefilter_line``; gather into the log-polar
  grids (3) vs ``mirror.interp``; row R2C (4) vs numpy's rfft; column pass
  (5) vs numpy's FFT x kernel spectrum x inverse FFT; all three FFT passes
  (6) vs ``mirror.lp_convolve``;
* end to end: slices of the 32-slice batch against ``mirror.fwd`` /
  ``mirror.adj`` (relative L2 <= 1e-5), and every slice of the batch against
  a one-slice handle (<= 1e-6: the same kernels without slice grouping);
* the same at N0 = 1024, 1024 angles (rows of h = 512, columns of 2048) and
  N0 = 2048, 2048 angles; a 32-slice batch at N0 = 512 against the mirror.
"""
import gc

import numpy as np
import pytest

from oracle import mirror
from lprec_amd import lpTransform
from lprec_amd.phantom import phantom_stack

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def make(N0, P0, ns, filt="none", cor=None, interp="cubic"):
    lp = lpTransform(N0, P0, ns, filt, N0 / 2 if cor is None else cor, interp)
    lp.precompute(1)
    lp.initcmem(1, 0)
    tab = mirror.Tables(lp.glpars, lp.fwdparsi, lp.fwdparamsf, lp.adjparsi, lp.adjparamsf)
    return lp, tab


def release():
    """Free device memory of handles the caller has dropped."""
    import torch
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def packed_rfft(F):
    """Half spectrum along theta of F [..., Nr, Nt] in the row pass's layout:
    [..., Nt/2, Nr] (rho contiguous), unnormalised, with the real theta-DC and
    theta-Nyquist terms packed as X0 + i X_{Nt/2} in frequency slot 0."""
    h = F.shape[-1] // 2
    X = np.fft.rfft(F, axis=-1)
    P = X[..., :h].copy()
    P[..., 0] = X[..., 0].real + 1j * X[..., h].real
    return np.swapaxes(P, -1, -2)


def column_pass(S, fZ, Nt):
    """FFT along rho of every half-spectrum column, times the kernel spectrum
    column, inverse FFT, with the 1 / (Ntheta Nrho) scale of the convolution
    (lpRgpu.cu:270-277): S [..., h, Nr] packed -> same layout."""
    h = S.shape[-2]
    fzT = np.asarray(fZ).T                         # [h + 1, Nr]
    out = np.fft.ifft(np.fft.fft(S, axis=-1) * fzT[:h], axis=-1) / Nt
    dc = np.fft.ifft(np.fft.fft(S[..., 0, :].real, axis=-1) * fzT[0], axis=-1) / Nt
    ny = np.fft.ifft(np.fft.fft(S[..., 0, :].imag, axis=-1) * fzT[h], axis=-1) / Nt
    # the C2R uses the real parts of the theta-DC and Nyquist columns
    out[..., 0, :] = dc.real + 1j * ny.real
    return out


def valid_mask(tab, adjoint):
    v = np.zeros(tab.Nrho * tab.Ntheta, bool)
    if adjoint:
        v[tab.adj["lpids"]] = True
        v[tab.adj["wids"]] = True
    else:
        v[tab.fwd["cids"]] = True
    return v.reshape(tab.Nrho, tab.Ntheta)


def gather_lp(tab, src, k, adjoint):
    """Stage 3 of one span from the device's prefiltered source src [z, H, W]."""
    grid = np.zeros((src.shape[0], tab.Nrho * tab.Ntheta))
    if adjoint:
        aj = tab.adj
        grid[:, aj["lpids"]] += mirror.interp(src, aj["lp2p2"][k], aj["lp2p1"][k], tab.cubic)
        grid[:, aj["wids"]] += mirror.interp(src, aj["lp2p2w"][k], aj["lp2p1w"][k], tab.cubic)
    else:
        fw = tab.fwd
        grid[:, fw["cids"]] = mirror.interp(src, fw["lp2C1"][k], fw["lp2C2"][k], tab.cubic)
        grid *= tab.erho.reshape(-1)
    return grid.reshape(src.shape[0], tab.Nrho, tab.Ntheta)


def stage_check(lp, tab, ns0, pick, adjoint, seed):
    """Run the 32-slice (ns0) call stage by stage and compare the picked slices."""
    import torch
    h = lp.clphandle[0]
    N, P, Nr, Nt = tab.N, tab.Nproj, tab.Nrho, tab.Ntheta
    N0, P0 = tab.N0, tab.Nproj // tab.os
    rng = np.random.default_rng(seed)
    if adjoint:
        x = rng.standard_normal((ns0, P0, N0)).astype(np.float32)
        out = torch.zeros((ns0, N0, N0), device="cuda")
        xin = torch.from_numpy(x).cuda()
        run = lambda: lp.adjp(out, xin, 0)  # noqa: E731
        srcbuf, rows = 1, P
    else:
        x = rng.standard_normal((ns0, N0, N0)).astype(np.float32)
        out = torch.zeros((ns0, P0, N0), device="cuda")
        xin = torch.from_numpy(x).cuda()
        run = lambda: lp.fwdp(out, xin, 0)  # noqa: E731
        srcbuf, rows = 0, N
    pick = list(pick)
    hh = Nt // 2
    try:
        h.debug_stop(2)
        run()
        src = h.debug_copy(srcbuf, (ns0, rows, N))[pick].astype(np.float64)
        h.debug_stop(3)
        run()
        F3 = h.debug_copy(2, (3, ns0, Nr, Nt))[:, pick].astype(np.float64)
        h.debug_stop(4)
        run()
        S4 = h.debug_copy(3, (3, ns0, hh, Nr), np.complex64)[:, pick].astype(np.complex128)
        h.debug_stop(5)
        run()
        S5 = h.debug_copy(3, (3, ns0, hh, Nr), np.complex64)[:, pick].astype(np.complex128)
        h.debug_stop(6)
        run()
        G6 = h.debug_copy(2, (3, ns0, Nr, Nt))[:, pick].astype(np.float64)
    finally:
        h.debug_stop(0)
    # stage 2: the device's prefiltered source vs the restated prep + prefilter
    if adjoint:
        d = mirror.sino_prepare(tab, x[pick])
    else:
        d = np.zeros((len(pick), N, N))
        o = N // 2 - N0 // 2
        d[:, o:o + N0, o:o + N0] = x[pick]
    e2 = rel(src, mirror.prefilter2d(d) if tab.cubic else d)
    valid = valid_mask(tab, adjoint)
    fz = tab.adj["fZ"] if adjoint else tab.fwd["fZ"]
    e3 = e4 = e5 = e6 = 0.0
    for k in range(3):
        ref3 = gather_lp(tab, src, k, adjoint)
        e3 = max(e3, rel(F3[k][:, valid], ref3[:, valid]))
        F3m = np.where(valid, F3[k], 0.0)
        e4 = max(e4, rel(S4[k], packed_rfft(F3m)))
        e5 = max(e5, rel(S5[k], column_pass(S4[k], fz, Nt)))
        e6 = max(e6, rel(G6[k], mirror.lp_convolve(F3m, fz)))
    return e2, e3, e4, e5, e6


@pytest.fixture(scope="module")
def headline(torch_cuda):
    print("headline handle: N0 2048, 1500 angles, 32 slices", flush=True)
    lp, tab = make(2048, 1500, 32)
    assert (tab.N, tab.Nrho, tab.Ntheta, tab.os) == (2048, 4096, 2048, 2)
    yield lp, tab
    del lp
    release()


@pytest.mark.parametrize("adjoint", [False, True], ids=["fwd", "adj"])
def test_headline_stages_match_restated_stages(headline, adjoint):
    lp, tab = headline
    print("headline stages", "adj" if adjoint else "fwd", flush=True)
    e = stage_check(lp, tab, 32, (0, 13, 31), adjoint, seed=11 + adjoint)
    print("stage errors: prefilter %.2e gather %.2e rowfft %.2e columns %.2e convolution %.2e" % e, flush=True)
    assert e[0] < 2e-6 and e[1] < 5e-6 and e[2] < 2e-6 and e[3] < 5e-6 and e[4] < 1e-5


@pytest.mark.parametrize("adjoint", [False, True], ids=["fwd", "adj"])
def test_radix16_inverse_rows_match_restated_stages(headline, adjoint):
    """FFT kernel family 1 (key 4): the radix-16 inverse row kernel
    (k_rows16_inv<4>) in place of the default 32-lane one (k_rows32_inv) --
    the whole convolution (stage 6) against numpy."""
    lp, tab = headline
    h = lp.clphandle[0]
    old = h.get_tuning(4)
    try:
        h.set_tuning(4, 1)
        print("headline stages, radix-16 inverse rows,", "adj" if adjoint else "fwd", flush=True)
        e = stage_check(lp, tab, 32, (0, 13, 31), adjoint, seed=21 + adjoint)
    finally:
        h.set_tuning(4, old)
    print("stage errors: prefilter %.2e gather %.2e rowfft %.2e columns %.2e convolution %.2e" % e, flush=True)
    assert e[2] < 2e-6 and e[3] < 5e-6 and e[4] < 1e-5


def test_row_kernel_families_agree(headline, torch_cuda):
    """The four operators with the 32-lane inverse row kernel (family 2, the
    default) against the radix-16 one (family 1) on the 32-slice batch
    (different FFT factorisations, so equal to rounding)."""
    torch = torch_cuda
    lp, tab = headline
    h = lp.clphandle[0]
    ns, N0, P0 = 32, 2048, 1500
    rng = np.random.default_rng(32)
    f = torch.from_numpy(rng.standard_normal((ns, N0, N0)).astype(np.float32)).cuda()
    R = torch.from_numpy(rng.standard_normal((ns, P0, N0)).astype(np.float32)).cuda()
    old = h.get_tuning(4)
    outs = []
    try:
        for fam in (1, 2):
            h.set_tuning(4, fam)
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
            del Ro, fo, fT, RT
    finally:
        h.set_tuning(4, old)
    worst = 0.0
    for name, a, b in zip(("fwd", "adj", "fwd_T", "adj_T"), outs[1], outs[0]):
        d = rel(a, b)
        print("%s: 32-lane vs radix-16 inverse row kernel %.2e" % (name, d), flush=True)
        worst = max(worst, d)
    release()
    assert worst <= 1e-6


def test_headline_batch_matches_mirror(headline):
    import torch
    lp, tab = headline
    ns, N0, P0 = 32, 2048, 1500
    f = phantom_stack(ns, N0, 20, seed=2048)
    R = np.random.default_rng(1500).standard_normal((ns, P0, N0)).astype(np.float32)
    Rout = torch.zeros((ns, P0, N0), device="cuda")
    fout = torch.zeros((ns, N0, N0), device="cuda")
    lp.fwdp(Rout, torch.from_numpy(f).cuda(), 0)
    lp.adjp(fout, torch.from_numpy(R).cuda(), 0)
    torch.cuda.synchronize()
    pick = [0, 1, 17, 31]
    ef = rel(Rout.cpu().numpy()[pick], mirror.fwd(tab, f[pick]))
    ea = rel(fout.cpu().numpy()[pick], mirror.adj(tab, R[pick]))
    print("headline 32-slice batch, slices %s: fwd %.2e adj %.2e" % (pick, ef, ea), flush=True)
    assert ef <= TOL and ea <= TOL


def test_headline_batch_matches_one_slice_handle(headline, torch_cuda):
    torch = torch_cuda
    lp, tab = headline
    ns, N0, P0 = 32, 2048, 1500
    rng = np.random.default_rng(77)
    f = torch.from_numpy(rng.standard_normal((ns, N0, N0)).astype(np.float32)).cuda()
    R = torch.from_numpy(rng.standard_normal((ns, P0, N0)).astype(np.float32)).cuda()
    Rb = torch.zeros((ns, P0, N0), device="cuda")
    fb = torch.zeros((ns, N0, N0), device="cuda")
    lp.fwdp(Rb, f, 0)
    lp.adjp(fb, R, 0)
    one, _ = make(N0, P0, 1)
    R1 = torch.zeros((1, P0, N0), device="cuda")
    f1 = torch.zeros((1, N0, N0), device="cuda")
    worst = 0.0
    for z in range(ns):
        one.fwdp(R1, f[z:z + 1].contiguous(), 0)
        one.adjp(f1, R[z:z + 1].contiguous(), 0)
        worst = max(worst, rel(R1.cpu().numpy()[0], Rb[z].cpu().numpy()),
                    rel(f1.cpu().numpy()[0], fb[z].cpu().numpy()))
    print("32-slice batch vs one-slice calls: worst relative difference %.2e" % worst, flush=True)
    del one
    release()
    assert worst <= 1e-6


@pytest.mark.parametrize("N0,P0,ns,pick", [(1024, 1024, 16, (0, 7, 15)), (2048, 2048, 4, (0, 3)),
                                            (4096, 1500, 2, (1,))],
                         ids=["N1024_P1024", "N2048_P2048", "N4096_P1500"])
def test_large_geometries_match_mirror(torch_cuda, N0, P0, ns, pick):
    torch = torch_cuda
    print("geometry", N0, P0, ns, flush=True)
    lp, tab = make(N0, P0, ns)
    f = phantom_stack(ns, N0, 20, seed=N0 + P0)
    R = np.random.default_rng(P0).standard_normal((ns, P0, N0)).astype(np.float32)
    Rout = torch.zeros((ns, P0, N0), device="cuda")
    fout = torch.zeros((ns, N0, N0), device="cuda")
    lp.fwdp(Rout, torch.from_numpy(f).cuda(), 0)
    lp.adjp(fout, torch.from_numpy(R).cuda(), 0)
    torch.cuda.synchronize()
    pick = list(pick)
    ef = rel(Rout.cpu().numpy()[pick], mirror.fwd(tab, f[pick]))
    ea = rel(fout.cpu().numpy()[pick], mirror.adj(tab, R[pick]))
    print("N0 %d P0 %d, slices %s: fwd %.2e adj %.2e" % (N0, P0, pick, ef, ea), flush=True)
    if N0 == 1024:
        e = stage_check(lp, tab, ns, (0, 15), True, seed=5)
        print("stage errors (adj): prefilter %.2e gather %.2e rowfft %.2e columns %.2e convolution %.2e" % e,
              flush=True)
        assert e[0] < 2e-6 and e[1] < 5e-6 and e[2] < 2e-6 and e[3] < 5e-6 and e[4] < 1e-5
    del lp
    release()
    assert ef <= TOL and ea <= TOL


def test_batch_of_32_slices_at_512_matches_mirror(torch_cuda):
    torch = torch_cuda
    N0, P0, ns = 512, 400, 32
    lp, tab = make(N0, P0, ns)
    f = phantom_stack(ns, N0, 12, seed=32)
    R = np.random.default_rng(33).standard_normal((ns, P0, N0)).astype(np.float32)
    Rout = torch.zeros((ns, P0, N0), device="cuda")
    fout = torch.zeros((ns, N0, N0), device="cuda")
    lp.fwdp(Rout, torch.from_numpy(f).cuda(), 0)
    lp.adjp(fout, torch.from_numpy(R).cuda(), 0)
    torch.cuda.synchronize()
    Rm, fm = mirror.fwd(tab, f), mirror.adj(tab, R)
    Rg, fg = Rout.cpu().numpy(), fout.cpu().numpy()
    ef = max(rel(Rg[z], Rm[z]) for z in range(ns))
    ea = max(rel(fg[z], fm[z]) for z in range(ns))
    print("N0 512, 32 slices: worst slice fwd %.2e adj %.2e" % (ef, ea), flush=True)
    del lp
    release()
    assert ef <= TOL and ea <= TOL


@pytest.mark.parametrize("N0,P0,ns,interp", [(2048, 1500, 32, "cubic"), (1024, 1024, 4, "cubic"),
                                             (512, 400, 8, "cubic"), (256, 384, 4, "linear")],
                         ids=["N2048_b32", "N1024", "N512", "N256_linear"])
def test_fused_polar_output_matches_two_kernel_path(headline, torch_cuda, N0, P0, ns, interp):
    """The forward operator's fused inverse row pass + polar gather (tuning
    key 6 = 1) against the separate kernels (key 6 = 0, the default): the same
    arithmetic on the same values (relative L2 <= 1e-6)."""
    torch = torch_cuda
    if N0 == 2048:
        lp, _ = headline
    else:
        lp, _ = make(N0, P0, ns, interp=interp)
    h = lp.clphandle[0]
    f = torch.from_numpy(np.random.default_rng(N0).standard_normal((ns, N0, N0)).astype(np.float32)).cuda()
    Rf = torch.zeros((ns, P0, N0), device="cuda")
    Ru = torch.full((ns, P0, N0), 3.0, device="cuda")
    try:
        h.set_tuning(6, 1)
        lp.fwdp(Rf, f, 0)
        h.set_tuning(6, 0)
        lp.fwdp(Ru, f, 0)
    finally:
        h.set_tuning(6, 0)
    torch.cuda.synchronize()
    a, b = Rf.cpu().numpy(), Ru.cpu().numpy()
    e = rel(a, b)
    print("N0 %d %s: fused vs two kernels relative %.2e, max abs %.2e" % (N0, interp, e, np.max(np.abs(a - b))),
          flush=True)
    assert e <= 1e-6
