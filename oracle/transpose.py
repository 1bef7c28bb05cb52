"""Exact transposes of the mirror operators (TEST INFRASTRUCTURE).

``fwd_T`` is the transpose of ``mirror.fwd`` and ``adj_T`` the transpose of
``mirror.adj``, stage by stage in reverse order:

* a gather ``y = interp(T, x, y)`` becomes a scatter-add of each sample into
  its 16 (4) B-spline taps with the same weights and periodic addressing;
* the log-polar convolution ``irfft2(rfft2(F) * fZ)`` becomes
  ``irfft2(rfft2(G) * conj(fZ))`` (the half-spectrum weights of rfft/irfft
  cancel, and the complex multiplier is conjugated);
* the B-spline prefilter (main_kernels.cuh:97-142) becomes its transposed
  recursion: a causal sweep ``s[m] = v[m] + z s[m-1]`` with outputs
  ``-z s[m]`` (``z/(z-1) s[L-1]`` at the end), then an anticausal sweep
  ``t[m] = u[m] + z t[m+1]`` with outputs ``lam t[m]``, plus
  ``lam z^(m+1) t[0]`` on the first 12 entries (the causal initialisation)
  and ``lam (1+z) t[0]`` at m = 0;
* pad / crop, angle stuffing / decimation and the edge replication swap into
  each other; the FBP filter is a symmetric Toeplitz matrix (a real even
  kernel), so it is its own transpose.

The reference has no transposes (its ``adj`` is an independently derived
operator, initsadj.py:136 vs initsfwd.py:86).  Pins: the dot-product test
<A x, y> = <x, A^T y> against ``mirror.fwd`` / ``mirror.adj`` in FP64, and the
per-stage dot tests in tests/test_transpose.py.
"""
import numpy as np
import scipy.fft

from .mirror import HORIZON, POLE, WORKERS, apply_filter, bspline_weights, prefilter_line


def interp_T(vals, x, y, H, W, cubic=True):
    """Transpose of mirror.interp: scatter vals[..., len(x)] into [..., H, W]."""
    vals = np.asarray(vals, dtype=np.float64)
    lead = vals.shape[:-1]
    v2 = vals.reshape(-1, vals.shape[-1])
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    ix = np.floor(x)
    iy = np.floor(y)
    fx = x - ix
    fy = y - iy
    ix = ix.astype(np.int64)
    iy = iy.astype(np.int64)
    if cubic:
        wx = bspline_weights(fx)
        wy = bspline_weights(fy)
        taps = [(((iy - 1 + b) % H) * W + (ix - 1 + a) % W, wy[b] * wx[a]) for b in range(4) for a in range(4)]
    else:
        wx = (1.0 - fx, fx)
        wy = (1.0 - fy, fy)
        taps = [(((iy + b) % H) * W + (ix + a) % W, wy[b] * wx[a]) for b in range(2) for a in range(2)]
    out = np.zeros((v2.shape[0], H * W))
    for s in range(v2.shape[0]):
        for idx, w in taps:
            out[s] += np.bincount(idx, weights=w * v2[s], minlength=H * W)
    return out.reshape(lead + (H, W))


def prefilter_line_T(v, axis):
    """Transpose of mirror.prefilter_line along ``axis``."""
    v = np.moveaxis(np.array(v, dtype=np.float64), axis, 0)
    L = v.shape[0]
    z = POLE
    lam = (1.0 - z) * (1.0 - 1.0 / z)
    hz = min(HORIZON, L)
    # transpose of the anticausal stage
    s = np.empty_like(v)
    s[0] = v[0]
    for m in range(1, L):
        s[m] = v[m] + z * s[m - 1]
    u = -z * s
    u[L - 1] = (z / (z - 1.0)) * s[L - 1]
    # transpose of the causal stage (with its 12-term initialisation)
    t = np.empty_like(u)
    t[L - 1] = u[L - 1]
    for m in range(L - 2, -1, -1):
        t[m] = u[m] + z * t[m + 1]
    out = lam * t
    out[0] = lam * (1.0 + z) * t[0]
    for m in range(1, hz):
        out[m] += lam * z ** (m + 1) * t[0]
    return np.moveaxis(out, 0, axis)


def prefilter2d_T(a):
    """Transpose of mirror.prefilter2d (x sweep, then y sweep): y^T, then x^T."""
    return prefilter_line_T(prefilter_line_T(a, -2), -1)


def lp_convolve_T(G, fZ):
    Nrho, Ntheta = G.shape[-2:]
    return scipy.fft.irfft2(scipy.fft.rfft2(G, workers=WORKERS) * np.conj(fZ), s=(Nrho, Ntheta), workers=WORKERS)


def fwd_T(tab, R):
    """Transpose of mirror.fwd: R [Ns0, P0, N0] -> [Ns0, N0, N0]."""
    R = np.asarray(R, dtype=np.float64)
    Ns0, P0 = R.shape[0], R.shape[1]
    N, N0, P, os_ = tab.N, tab.N0, tab.Nproj, tab.os
    Nt, Nr = tab.Ntheta, tab.Nrho
    c0 = N // 2 - tab.cor
    dR = np.zeros((Ns0, P, N))
    dR[:, ::os_][:, :P0, c0:c0 + N0] = R
    dR = dR.reshape(Ns0, P * N)
    fw = tab.fwd
    df = np.zeros((Ns0, N, N))
    for k in range(tab.Nspan):
        G = interp_T(dR[:, fw["pids"][k]], fw["p2lp1"][k], fw["p2lp2"][k], Nr, Nt, tab.cubic)
        F = lp_convolve_T(G, fw["fZ"]) * tab.erho
        df += interp_T(F.reshape(Ns0, -1)[:, fw["cids"]], fw["lp2C1"][k], fw["lp2C2"][k], N, N, tab.cubic)
    if tab.cubic:
        df = prefilter2d_T(df)
    o = N // 2 - N0 // 2
    return df[:, o:o + N0, o:o + N0]


def sino_prepare_T(tab, dR, P0):
    """Transpose of mirror.sino_prepare: [Ns0, P, N] -> [Ns0, P0, N0]."""
    dR = np.array(dR, dtype=np.float64)
    N, N0, os_ = tab.N, tab.N0, tab.os
    lb = N // 2 - tab.cor
    rb = N // 2 + N0 - tab.cor - 1
    if tab.adj["filter"] is not None:
        dR = apply_filter(dR, tab.adj["filter"])
    dR[:, :, lb] += dR[:, :, :lb].sum(axis=-1)
    dR[:, :, rb] += dR[:, :, rb + 1:].sum(axis=-1)
    return dR[:, ::os_][:, :P0, lb:lb + N0]


def adj_T(tab, f, P0):
    """Transpose of mirror.adj: f [Ns0, N0, N0] -> [Ns0, P0, N0]."""
    f = np.asarray(f, dtype=np.float64)
    Ns0 = f.shape[0]
    N, N0, P = tab.N, tab.N0, tab.Nproj
    Nt, Nr = tab.Ntheta, tab.Nrho
    o = N // 2 - N0 // 2
    df = np.zeros((Ns0, N, N))
    df[:, o:o + N0, o:o + N0] = f
    df = df.reshape(Ns0, N * N)
    aj = tab.adj
    dR = np.zeros((Ns0, P, N))
    for k in range(tab.Nspan):
        G = interp_T(df[:, aj["cids"]], aj["C2lp1"][k], aj["C2lp2"][k], Nr, Nt, tab.cubic)
        F = lp_convolve_T(G, aj["fZ"]).reshape(Ns0, -1)
        dR += interp_T(F[:, aj["lpids"]], aj["lp2p2"][k], aj["lp2p1"][k], P, N, tab.cubic)
        dR += interp_T(F[:, aj["wids"]], aj["lp2p2w"][k], aj["lp2p1w"][k], P, N, tab.cubic)
    if tab.cubic:
        dR = prefilter2d_T(dR)
    return sino_prepare_T(tab, dR, P0)


__all__ = ["interp_T", "prefilter_line_T", "prefilter2d_T", "lp_convolve_T", "fwd_T", "adj_T", "sino_prepare_T",
           "prefilter_line"]
