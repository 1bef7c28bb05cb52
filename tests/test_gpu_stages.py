"""Stage-by-stage GPU parity against the FP64 mirror's intermediates.

Uses lp_debug_stop / lp_debug_copy: the pipeline returns after stage k and an
internal buffer is copied out.  Stages: 1 row sweep (placement / stuffing /
edge padding / FBP filter + x-direction prefilter), 2 column prefilter,
3 gather into the log-polar grids (x e^rho for fwd), 6 convolved log-polar
grids.  Geometries include a partial last prefilter chunk (Nproj = 300), a
centre offset and a filter.
"""
import numpy as np
import pytest

from oracle import mirror
from lprec_amd import lpTransform

CASES = [
    # N0, P0, filter, cor, interp
    (256, 300, "none", 128, "cubic"),
    (250, 300, "ramp", 127.3, "cubic"),
    (256, 384, "shepp-logan", 128, "cubic"),
    (240, 200, "none", 120, "linear"),
    (256, 150, "none", 128, "cubic"),         # os = 3: stuffed angle rows
    (256, 180, "ramp", 128, "cubic"),         # os = 2 with the paired filter path
]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def mirror_stages(tab, x, adjoint):
    N, Nr, Nt = tab.N, tab.Nrho, tab.Ntheta
    if adjoint:
        d = mirror.sino_prepare(tab, x)
    else:
        d = np.zeros((x.shape[0], N, N))
        o = N // 2 - tab.N0 // 2
        d[:, o:o + tab.N0, o:o + tab.N0] = x
    s1 = mirror.prefilter_line(d, -1) if tab.cubic else d
    s2 = mirror.prefilter_line(s1, -2) if tab.cubic else d
    ns = x.shape[0]
    F = np.zeros((3, ns, Nr * Nt))
    valid = np.zeros((3, Nr * Nt), bool)
    for k in range(3):
        if adjoint:
            aj = tab.adj
            F[k][:, aj["lpids"]] += mirror.interp(s2, aj["lp2p2"][k], aj["lp2p1"][k], tab.cubic)
            F[k][:, aj["wids"]] += mirror.interp(s2, aj["lp2p2w"][k], aj["lp2p1w"][k], tab.cubic)
            valid[k, aj["lpids"]] = valid[k, aj["wids"]] = True
        else:
            fw = tab.fwd
            F[k][:, fw["cids"]] = mirror.interp(s2, fw["lp2C1"][k], fw["lp2C2"][k], tab.cubic)
            F[k] *= tab.erho.reshape(-1)
            valid[k, fw["cids"]] = True
    fz = tab.adj["fZ"] if adjoint else tab.fwd["fZ"]
    G = np.stack([mirror.lp_convolve(F[k].reshape(ns, Nr, Nt), fz) for k in range(3)])
    return s1, s2, F, valid, G


@pytest.mark.gpu
@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("N0,P0,filt,cor,interp", CASES)
def test_pipeline_stages_match_mirror(torch_cuda, N0, P0, filt, cor, interp, adjoint):
    torch = torch_cuda
    ns = 2
    print("case", N0, P0, filt, cor, interp, "adj" if adjoint else "fwd", flush=True)
    lp = lpTransform(N0, P0, ns, filt, cor, interp)
    lp.precompute(1)
    lp.initcmem(1, 0)
    h = lp.clphandle[0]
    tab = mirror.Tables(lp.glpars, lp.fwdparsi, lp.fwdparamsf, lp.adjparsi, lp.adjparamsf)
    N, P, Nr, Nt = tab.N, tab.Nproj, tab.Nrho, tab.Ntheta
    rng = np.random.default_rng(N0 + P0)
    if adjoint:
        x = rng.standard_normal((ns, P0, N0)).astype(np.float32)
        xin = torch.from_numpy(x).cuda()
        out = torch.zeros((ns, N0, N0), device="cuda")
        run = lambda: lp.adjp(out, xin, 0)  # noqa: E731
        rows, buf = P, 1
    else:
        x = rng.standard_normal((ns, N0, N0)).astype(np.float32)
        xin = torch.from_numpy(x).cuda()
        out = torch.zeros((ns, P0, N0), device="cuda")
        run = lambda: lp.fwdp(out, xin, 0)  # noqa: E731
        rows, buf = N, 0
    s1, s2, F, valid, G = mirror_stages(tab, x, adjoint)
    try:
        h.debug_stop(1)
        run()
        a1 = h.debug_copy(4 if tab.cubic else buf, (ns, rows, N))
        h.debug_stop(2)
        run()
        a2 = h.debug_copy(buf, (ns, rows, N))
        h.debug_stop(3)
        run()
        a3 = h.debug_copy(2, (3, ns, Nr * Nt))
        h.debug_stop(6)
        run()
        a6 = h.debug_copy(2, (3, ns, Nr, Nt))
    finally:
        h.debug_stop(0)
    if adjoint and tab.cubic:
        # the stuffed (zero) angle rows of the row-sweep buffer are not written:
        # the column sweep treats them as zero without reading them
        keep = np.arange(rows) % tab.os == 0
        a1, s1 = a1[:, keep], s1[:, keep]
    e1, e2 = rel(a1, s1), rel(a2, s2)
    m = np.broadcast_to(valid[:, None, :], a3.shape)
    e3 = rel(a3[m], F[m])
    e6 = rel(a6, G)
    print("stage errors %.2e %.2e %.2e %.2e" % (e1, e2, e3, e6), flush=True)
    assert e1 < 2e-6 and e2 < 2e-6 and e3 < 5e-6 and e6 < 1e-5
