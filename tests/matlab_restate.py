"""FP64 numpy restatement of the reference's MATLAB fixture generators
(TEST INFRASTRUCTURE, used by tests/test_oracle_filters.py).

The reference's fixtures tests/data/f, R (and fbub, Rbub) were written by
tests/matlab_files/make_phantom.m:1-15 (make_bubles.m:1-19):

* ``filtered_sinogram``  apply_filter_exact.m:1-41 with fRphantom.m:1-38: the
  Fourier transform of each ellipse's analytic projection, summed over a
  16N-point detector grid (the MATLAB ``ep * p`` sum), shifted by half a
  detector cell for even N, multiplied by the centred filter ``wfa``, inverse
  FFT, the central N cells kept.  MATLAB's ``'`` conjugate-transposes, so the
  transform is taken with exp(+2 pi i rho s) and the shift with
  exp(+2 pi i delta s).
* ``filtered_image``  apply_filter_2d_exact.m:1-66 with fphantom.m:1-24: the
  analytic 2-D Fourier transform of each ellipse (J1 Bessel), the radial
  window on the 4x-oversampled frequency grid, inverse 2-D FFT, central N x N.
* ``SHEPP_LOGAN``  MATLAB's phantom(N) default table (modified Shepp-Logan),
  the ``ellipse`` that make_phantom.m:4 passes on; ``BUBBLES`` the table of
  make_bubles.m:4-7.

The filter array ``wfa`` is supplied by the caller (the oracle's
take_filter), so reproducing the fixtures checks the oracle's filter
against data MATLAB produced.
"""
import numpy as np

# A, a, b, x0, y0, phi (degrees)
SHEPP_LOGAN = np.array([
    [1.0, .69, .92, 0.0, 0.0, 0.0],
    [-.8, .6624, .8740, 0.0, -.0184, 0.0],
    [-.2, .1100, .3100, .22, 0.0, -18.0],
    [-.2, .1600, .4100, -.22, 0.0, 18.0],
    [.1, .2100, .2500, 0.0, .35, 0.0],
    [.1, .0460, .0460, 0.0, .1, 0.0],
    [.1, .0460, .0460, 0.0, -.1, 0.0],
    [.1, .0460, .0230, -.08, -.605, 0.0],
    [.1, .0230, .0230, 0.0, -.606, 0.0],
    [.1, .0230, .0460, .06, -.605, 0.0],
])

BUBBLES = np.array([
    [1.0, .05, .05, 0.0, 0.0, 0.0],
    [1.0, .04, .04, -0.5, -0.5, 0.0],
    [1.0, .03, .03, -.3, -.3, 0.0],
])

OS = 4


def filtered_sinogram(N, Ntheta, ellipses, wfa):
    """-> h [Ntheta, N]: apply_filter_exact(Ntheta, N, filter, ellipse) with
    the centred length-4N filter ``wfa`` (all ones: the unfiltered data)."""
    Nse = OS * N
    Nrho = 4 * Nse
    s = np.arange(-(Nse // 2), Nse // 2) / 2 / OS
    rho = np.arange(-(Nrho // 2), Nrho // 2) / Nrho * 2
    th = np.arange(Ntheta) * np.pi / Ntheta
    p = np.zeros((Nrho, Ntheta))
    for A, a, b, x0, y0, phd in ellipses:
        phi = -phd * np.pi / 180
        rho0 = np.sqrt(a * a * np.cos(th - phi) ** 2 + b * b * np.sin(th - phi) ** 2)
        d = rho[:, None] + x0 * np.cos(th)[None] - y0 * np.sin(th)[None]
        inside = np.abs(d) <= rho0[None]
        chord = 2 * A * a * b * np.sqrt(np.maximum(rho0[None] ** 2 - d * d, 0.0)) / rho0[None] ** 2
        p += np.where(inside, chord, 0.0)
    G = (np.exp(2j * np.pi * s[:, None] * rho[None, :]) @ p) * (rho[1] - rho[0]) * Nse / 2 / OS
    if N % 2 == 0:
        G *= np.exp(2j * np.pi * (4 * OS / Nrho) * s)[:, None]
    g = np.fft.fftshift(np.fft.ifft(np.fft.ifftshift(np.asarray(wfa, np.float64)[:, None] * G, axes=0), axis=0),
                        axes=0)
    g = g[Nse // 2 + np.arange(-(N // 2), (N + 1) // 2), :]
    return np.real(g).T


def fixture_sinogram_scale(N, Ntheta):
    """make_phantom.m:14: the stored sinogram is flipud(h') * 2N sqrt(2) sqrt(N) / sqrt(Ntheta)."""
    return 2 * N * np.sqrt(2) * np.sqrt(N) / np.sqrt(Ntheta)


def image_spectrum(N, ellipses):
    """sum over ellipses of fphantom(os, Ne, ellipse, doshift)' (conjugate
    transpose, apply_filter_2d_exact.m:16-18): [4N, 4N] complex."""
    from scipy.special import j1
    Ne = OS * N
    kk = np.arange(-(Ne // 2), Ne // 2 + 1) / 2 / OS
    k1, k2 = np.meshgrid(kk, kk, indexing="ij")
    F = np.zeros((Ne, Ne), complex)
    for A, a, b, x0, y0, phd in ellipses:
        phi = phd * np.pi / 180
        k1h = np.cos(phi) * k1 - np.sin(phi) * k2
        k2h = np.sin(phi) * k1 + np.cos(phi) * k2
        K = np.sqrt((a * k1h) ** 2 + (b * k2h) ** 2)
        with np.errstate(invalid="ignore", divide="ignore"):
            fh = A * a * b * np.pi * j1(2 * np.pi * K) / (np.pi * K) * np.exp(-2j * np.pi * (k1 * (-x0) + k2 * y0))
        fh[Ne // 2, Ne // 2] = A * a * b * np.pi
        fh *= (Ne / OS / 2) ** 2
        if N % 2 == 0:
            fh *= np.exp(-2j * np.pi * (OS / Ne * k1 + OS / Ne * k2))
        F += np.conj(fh[:-1, :-1]).T
    return F


def window_2d(name, t, d=0.5):
    """The radial windows of apply_filter_2d_exact.m:24-38 (parzen, which the
    MATLAB file lacks, as the reference's initsadj.py:164-165 defines it)."""
    if name == "hamming":
        return (.54 + .46 * np.cos(np.pi * t / d)) * (t / d <= 1)
    if name == "hann":
        return (1 + np.cos(np.pi * t / d)) / 2 * (t / d <= 1)
    if name == "cosine":
        return np.cos(np.pi * t / (2 * d)) * (t / d <= 1)
    if name == "cosine2":
        return np.cos(np.pi * t / (2 * d)) ** 2 * (t / d <= 1)
    if name == "shepp-logan":
        return np.sinc(t / (2 * d)) * (t / d <= 2 * np.sqrt(2))
    if name == "parzen":
        return (1 - t / d) ** 3 * (t / d <= 1)
    raise ValueError(name)


def filtered_image(N, F, name):
    """apply_filter_2d_exact(f, name, ellipse) from the spectrum F of
    ``image_spectrum``: [N, N] in MATLAB's orientation."""
    Ne = F.shape[0]
    tt = np.arange(0, Ne // 2 + 1) / Ne
    t1, t2 = np.meshgrid(tt, tt, indexing="ij")
    w = window_2d(name, np.sqrt(t1 ** 2 + t2 ** 2))
    w = w * (w >= 0)
    top = np.hstack([np.flipud(np.fliplr(w[1:, :])), np.flipud(w[1:, 1:])])
    bot = np.hstack([np.fliplr(w), w[:, 1:]])
    W = np.vstack([top, bot])[:-1, :-1]
    g = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(W * F)))
    sl = Ne // 2 + np.arange(-(N // 2), (N + 1) // 2)
    return np.real(g[np.ix_(sl, sl)])
