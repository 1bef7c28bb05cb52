"""FP64 numpy mirror of the reference device pipeline (TEST INFRASTRUCTURE).

Consumes the reference wire arrays (glpars, fwdparsi/fwdparamsf,
adjparsi/adjparamsf; layout params.cu:40-145) exactly as the reference's C++
class does, and restates:

* ``fwd``   lpRgpu::execFwdManyPtr + execFwd   (src/cuda/lpRgpu.cu:184-193, 243-286)
* ``adj``   lpRgpu::execAdjManyPtr + execAdj + padding + applyFilter
            (src/cuda/lpRgpu.cu:196-216, 289-382)
* ``prefilter_line``  ConvertToInterpolationCoefficients and its causal /
            anticausal initialisations (src/include/main_kernels.cuh:97-142)
* ``interp``  the ``interp`` kernel with cubicTex2D / linearTex2D
            (main_kernels.cuh:9-84): the sample at table coordinate (x, y)
            is the exact 16-tap cubic B-spline sum sum_ab w_a(fx) w_b(fy)
            T[(iy-1+b) mod H, (ix-1+a) mod W] with ix = floor(x), fx = x - ix
            (what the 4 bilinear fetches at h0/h1 compute, with exact
            instead of 8-bit-fraction weights), and periodic addressing
            (cudaAddressModeWrap, lpRgpu.cu:22-23, 77-78, 114-115).
            ``interp_taps`` is the numpy form; oracle/interp.c the same sum
            in C (OpenMP), used when built (``python -m oracle.build``).

Pins: ``prefilter_line`` against scipy.ndimage.spline_filter1d(mode='reflect');
``interp`` / ``interp_taps`` against scipy.ndimage.map_coordinates(order=3, mode='grid-wrap',
prefilter=False) and against brute-force 16-tap sums; the FFT convolution
against a direct circular convolution on small grids; the whole pipeline
against the reference fixtures (tests/golden/f, R) at the reference's
tolerances and against the analytic ellipse Radon transform.
"""
import os

import numpy as np
import scipy.fft

WORKERS = len(os.sched_getaffinity(0))

POLE = np.sqrt(3.0) - 2.0  # main_kernels.cuh:2
HORIZON = 12               # main_kernels.cuh:99


class Tables:
    """Parsed wire arrays (params.cu:40-145)."""

    def __init__(self, glpars, fwdparsi=None, fwdparamsf=None, adjparsi=None, adjparamsf=None):
        g = glpars
        self.N, self.N0, self.Ntheta, self.Nrho = int(g[0]), int(g[1]), int(g[2]), int(g[3])
        self.Nspan, self.Nproj, self.Nslices = int(g[4]), int(g[5]), int(g[6])
        self.cor, self.os, self.cubic = int(g[7]), int(g[8]), bool(int(g[9]))
        self.erho = np.asarray(g[10:10 + self.Nrho * self.Ntheta], dtype=np.float64).reshape(self.Nrho, self.Ntheta)
        H = self.Ntheta // 2 + 1
        self.fwd = self.adj = None
        if fwdparsi is not None:
            pi, pf = np.asarray(fwdparsi), np.asarray(fwdparamsf, dtype=np.float64)
            si = sf = 0
            Npids = [int(v) for v in pi[0:self.Nspan]]
            si = self.Nspan
            pids = []
            for k in range(self.Nspan):
                pids.append(pi[si:si + Npids[k]].astype(np.int64))
                si += Npids[k]
            Nc = int(pi[si])
            si += 1
            cids = pi[si:si + Nc].astype(np.int64)
            lp2C1, lp2C2, p2lp1, p2lp2 = [], [], [], []
            for dst, n in ((lp2C1, [Nc] * 3), (lp2C2, [Nc] * 3), (p2lp1, Npids), (p2lp2, Npids)):
                for k in range(self.Nspan):
                    dst.append(pf[sf:sf + n[k]])
                    sf += n[k]
            fz = pf[sf:sf + 2 * H * self.Nrho]
            self.fwd = dict(pids=pids, cids=cids, lp2C1=lp2C1, lp2C2=lp2C2, p2lp1=p2lp1, p2lp2=p2lp2,
                            fZ=(fz[0::2] + 1j * fz[1::2]).reshape(self.Nrho, H))
        if adjparsi is not None:
            pi, pf = np.asarray(adjparsi), np.asarray(adjparamsf, dtype=np.float64)
            Nc, Nl, Nw = int(pi[0]), int(pi[1]), int(pi[2])
            si = 3
            cids = pi[si:si + Nc].astype(np.int64); si += Nc
            lpids = pi[si:si + Nl].astype(np.int64); si += Nl
            wids = pi[si:si + Nw].astype(np.int64); si += Nw
            flag = int(pi[si])
            sf = 0
            out = {}
            for name, n in (("C2lp1", Nc), ("C2lp2", Nc), ("lp2p1", Nl), ("lp2p2", Nl),
                            ("lp2p1w", Nw), ("lp2p2w", Nw)):
                out[name] = []
                for k in range(self.Nspan):
                    out[name].append(pf[sf:sf + n])
                    sf += n
            fz = pf[sf:sf + 2 * H * self.Nrho]
            sf += 2 * H * self.Nrho
            out["fZ"] = (fz[0::2] + 1j * fz[1::2]).reshape(self.Nrho, H)
            out["filter"] = pf[sf:sf + 4 * self.N] if flag == 1 else None
            out.update(cids=cids, lpids=lpids, wids=wids)
            self.adj = out


def bspline_weights(f):
    """main_kernels.cuh:9-19 in FP64; f is the fractional coordinate."""
    one = 1.0 - f
    return (one ** 3 / 6.0,
            2.0 / 3.0 - 0.5 * f * f * (2.0 - f),
            2.0 / 3.0 - 0.5 * one * one * (2.0 - one),
            f ** 3 / 6.0)


def interp_taps(T, x, y, cubic=True):
    """Sample the periodic grid T[..., H, W] at table coordinates (x, y):
    the numpy restatement of the 16-tap (4-tap linear) sum.

    Leading axes of T are slices; returns [..., len(x)].
    """
    H, W = T.shape[-2:]
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    ix = np.floor(x)
    iy = np.floor(y)
    fx = x - ix
    fy = y - iy
    ix = ix.astype(np.int64)
    iy = iy.astype(np.int64)
    flat = T.reshape(T.shape[:-2] + (H * W,))
    out = np.zeros(T.shape[:-2] + (len(x),))
    if cubic:
        wx = bspline_weights(fx)
        wy = bspline_weights(fy)
        for b in range(4):
            rows = ((iy - 1 + b) % H) * W
            acc = np.zeros_like(out)
            for a in range(4):
                acc += wx[a] * flat[..., rows + (ix - 1 + a) % W]
            out += wy[b] * acc
    else:
        for b, wyb in ((0, 1.0 - fy), (1, fy)):
            rows = ((iy + b) % H) * W
            out += wyb * ((1.0 - fx) * flat[..., rows + ix % W] + fx * flat[..., rows + (ix + 1) % W])
    return out


_CLIB = None


def _clib():
    """oracle/liboracle.so (oracle/interp.c, built by oracle/build.py), or None."""
    global _CLIB
    if _CLIB is None:
        import ctypes
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle.so")
        try:
            lib = ctypes.CDLL(path)
            d = ctypes.POINTER(ctypes.c_double)
            i64 = ctypes.c_int64
            lib.oracle_interp.argtypes = [d, i64, i64, i64, d, d, i64, ctypes.c_int, d]
            lib.oracle_interp.restype = None
            _CLIB = lib
        except OSError:
            _CLIB = False
    return _CLIB or None


def interp(T, x, y, cubic=True):
    """Sample the periodic grid T[..., H, W] at table coordinates (x, y).

    Leading axes of T are slices; returns [..., len(x)].  Runs the C
    restatement (oracle/interp.c) when it is built, else ``interp_taps``;
    the two agree to 1e-12 (tests/test_oracle.py).
    """
    lib = _clib()
    if lib is None:
        return interp_taps(T, x, y, cubic)
    import ctypes
    H, W = T.shape[-2:]
    lead = T.shape[:-2]
    Tc = np.ascontiguousarray(T, dtype=np.float64).reshape(-1, H, W)
    xc = np.ascontiguousarray(x, dtype=np.float64).ravel()
    yc = np.ascontiguousarray(y, dtype=np.float64).ravel()
    n = xc.size
    out = np.empty((Tc.shape[0], n))
    d = ctypes.POINTER(ctypes.c_double)
    lib.oracle_interp(Tc.ctypes.data_as(d), Tc.shape[0], H, W, xc.ctypes.data_as(d), yc.ctypes.data_as(d), n,
                      1 if cubic else 0, out.ctypes.data_as(d))
    return out.reshape(lead + (n,))


def prefilter_line(c, axis):
    """ConvertToInterpolationCoefficients (main_kernels.cuh:121-142) along ``axis``,
    with the 12-term causal init (:97-111) and the anticausal init (:114-118)."""
    c = np.moveaxis(np.array(c, dtype=np.float64), axis, 0)
    L = c.shape[0]
    z = POLE
    lam = (1.0 - z) * (1.0 - 1.0 / z)
    hz = min(HORIZON, L)
    s = c[0].copy()
    zn = z
    for n in range(hz):
        s += zn * c[n]
        zn *= z
    c[0] = lam * s
    for n in range(1, L):
        c[n] = lam * c[n] + z * c[n - 1]
    c[L - 1] = (z / (z - 1.0)) * c[L - 1]
    for n in range(L - 2, -1, -1):
        c[n] = z * (c[n + 1] - c[n])
    return np.moveaxis(c, 0, axis)


def prefilter2d(a):
    """prefilter2D (lpRgpu.cu:219-240): along rows (x) first, then columns (y)."""
    return prefilter_line(prefilter_line(a, -1), -2)


def lp_convolve(F, fZ):
    """cuFFT R2C -> mul(1/(Ntheta Nrho), fZ) -> C2R (lpRgpu.cu:270-277, simple_kernels.cuh:2-12).

    numpy's irfft2 already divides by Nrho*Ntheta, so the explicit scale is
    folded into it."""
    Nrho, Ntheta = F.shape[-2:]
    return scipy.fft.irfft2(scipy.fft.rfft2(F, workers=WORKERS) * fZ, s=(Nrho, Ntheta), workers=WORKERS)


def fwd(tab, f):
    """execFwdManyPtr: f [Ns0, N0, N0] -> R [Ns0, P0, N0] (lpRgpu.cu:184-193)."""
    f = np.asarray(f, dtype=np.float64)
    Ns0 = f.shape[0]
    N, N0, P, os_ = tab.N, tab.N0, tab.Nproj, tab.os
    Nt, Nr = tab.Ntheta, tab.Nrho
    df = np.zeros((Ns0, N, N))
    o = N // 2 - N0 // 2
    df[:, o:o + N0, o:o + N0] = f
    if tab.cubic:
        df = prefilter2d(df)
    fw = tab.fwd
    dR = np.zeros((Ns0, P * N))
    for k in range(tab.Nspan):
        dfl = np.zeros((Ns0, Nr * Nt))
        dfl[:, fw["cids"]] += interp(df, fw["lp2C1"][k], fw["lp2C2"][k], tab.cubic)
        dfl = dfl.reshape(Ns0, Nr, Nt) * tab.erho
        G = lp_convolve(dfl, fw["fZ"])
        dR[:, fw["pids"][k]] += interp(G, fw["p2lp1"][k], fw["p2lp2"][k], tab.cubic)
    dR = dR.reshape(Ns0, P, N)
    # copy3Dstep (keep every os-th angle) + copy3Dshifted (crop at N/2 - cor)
    c0 = N // 2 - tab.cor
    return dR[:, ::os_, c0:c0 + N0][:, :P // os_]


def sino_prepare(tab, R):
    """Steps of execAdjManyPtr before execAdj: place at column N/2 - cor,
    zero-stuff angles by os (stepangles), edge-replicate (padker), FBP filter
    (applyFilter).  R [Ns0, P0, N0] -> [Ns0, P, N]."""
    R = np.asarray(R, dtype=np.float64)
    Ns0, P0 = R.shape[0], R.shape[1]
    N, N0, P, os_ = tab.N, tab.N0, tab.Nproj, tab.os
    lb = N // 2 - tab.cor
    rb = N // 2 + N0 - tab.cor - 1
    dR = np.zeros((Ns0, P, N))
    dR[:, ::os_][:, :P0, lb:lb + N0] = R
    dR[:, :, :lb] = dR[:, :, lb:lb + 1]
    dR[:, :, rb + 1:] = dR[:, :, rb:rb + 1]
    filt = tab.adj["filter"]
    if filt is not None:
        dR = apply_filter(dR, filt)
    return dR


def apply_filter(dR, filt):
    """applyFilter (lpRgpu.cu:333-382): each row is zero-padded to 4N at
    offset 2N - N/2 (copyc), multiplied by (-1)^x before and after each C2C
    (fftshift), by the centred filter in frequency (mulfilter), cropped
    (copycback) and scaled by 1/(4N) (mulconst).  The four modulations make
    it a circular convolution of the padded row with ifftshift(filter)."""
    N = dR.shape[-1]
    M = 4 * N
    off = M // 2 - N // 2
    pad = np.zeros(dR.shape[:-1] + (M,))
    pad[..., off:off + N] = dR
    sgn = (-1.0) ** np.arange(M)
    X = np.fft.fft(pad * sgn) * sgn
    X *= np.asarray(filt, dtype=np.float64)
    y = np.fft.ifft(X * sgn) * M * sgn   # cuFFT inverse is unnormalised
    return y.real[..., off:off + N] / M


def adj(tab, R):
    """execAdjManyPtr: R [Ns0, P0, N0] -> f [Ns0, N0, N0] (lpRgpu.cu:196-207, 289-330)."""
    dR = sino_prepare(tab, R)
    Ns0 = dR.shape[0]
    N, N0 = tab.N, tab.N0
    Nt, Nr = tab.Ntheta, tab.Nrho
    if tab.cubic:
        dR = prefilter2d(dR)
    aj = tab.adj
    df = np.zeros((Ns0, N * N))
    for k in range(tab.Nspan):
        dfl = np.zeros((Ns0, Nr * Nt))
        dfl[:, aj["lpids"]] += interp(dR, aj["lp2p2"][k], aj["lp2p1"][k], tab.cubic)
        dfl[:, aj["wids"]] += interp(dR, aj["lp2p2w"][k], aj["lp2p1w"][k], tab.cubic)
        G = lp_convolve(dfl.reshape(Ns0, Nr, Nt), aj["fZ"])
        df[:, aj["cids"]] += interp(G, aj["C2lp1"][k], aj["C2lp2"][k], tab.cubic)
    o = N // 2 - N0 // 2
    return df.reshape(Ns0, N, N)[:, o:o + N0, o:o + N0]
