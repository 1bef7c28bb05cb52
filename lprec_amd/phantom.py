"""Seeded random-ellipse phantoms (synthetic inputs for tests and benchmarks).

An ellipse is (A, a, b, x0, y0, psi): intensity A, semi-axes a, b, centre
(x0, y0) in the unit-disk frame, rotation psi.  Images are sampled at the
pixel centres of linspace(-1, 1, N) x linspace(-1, 1, N) with x along columns
and y along rows, and masked to the unit disk (the log-polar operators only
see |x| <= 1, initsfwd.py:30 and initsadj.py:29).
"""
import numpy as np


def random_ellipses(n, rng, inside=False):
    """n ellipses: centres U(-0.6, 0.6)^2, semi-axes U(0.05, 0.4)^2, angle U(0, pi),
    intensity U(-0.3, 1).  ``inside=True`` redraws ellipses that leave the
    unit disk (needed when comparing with an analytic Radon transform)."""
    out = []
    while len(out) < n:
        A = rng.uniform(-0.3, 1.0)
        a, b = rng.uniform(0.05, 0.4, 2)
        x0, y0 = rng.uniform(-0.6, 0.6, 2)
        psi = rng.uniform(0, np.pi)
        if inside and np.hypot(x0, y0) + max(a, b) > 0.98:
            continue
        out.append((A, a, b, x0, y0, psi))
    return np.array(out)


def rasterize(ell, N, lin=None):
    """float64 [N, N] image of the ellipse set, masked to the unit disk, at
    pixel coordinates ``lin`` (default linspace(-1, 1, N), increasing).
    Each ellipse is evaluated on the pixels of its bounding box only."""
    lin = np.linspace(-1, 1, N) if lin is None else np.asarray(lin, np.float64)
    img = np.zeros((N, N))
    for A, a, b, x0, y0, psi in ell:
        c, s = np.cos(psi), np.sin(psi)
        # half extents of the rotated ellipse along x and y, plus a margin
        ex = np.sqrt((a * c) ** 2 + (b * s) ** 2) * 1.0001 + 1e-9
        ey = np.sqrt((a * s) ** 2 + (b * c) ** 2) * 1.0001 + 1e-9
        i0, i1 = np.searchsorted(lin, [x0 - ex, x0 + ex])
        j0, j1 = np.searchsorted(lin, [y0 - ey, y0 + ey])
        if i1 <= i0 or j1 <= j0:
            continue
        X, Y = np.meshgrid(lin[i0:i1], lin[j0:j1])
        u = (X - x0) * c + (Y - y0) * s
        v = -(X - x0) * s + (Y - y0) * c
        img[j0:j1, i0:i1][(u / a) ** 2 + (v / b) ** 2 <= 1] += A
    X, Y = np.meshgrid(lin, lin)
    img[X ** 2 + Y ** 2 > 1] = 0
    return img


def phantom_stack(nslices, N, n_ellipses, seed, inside=False):
    """float32 [nslices, N, N]: one independent seeded phantom per slice."""
    rng = np.random.default_rng(seed)
    return np.stack([rasterize(random_ellipses(n_ellipses, rng, inside), N)
                     for _ in range(nslices)]).astype(np.float32)


def radon_scale(N, P0):
    """Factor between lpTransform.fwdp output and line integrals in unit-disk
    lengths: the fwd kernel constant sqrt(N os / P) pi / 4 / aR / sqrt(2)
    (initsfwd.py:57-58) times aR, with P = os P0 (N = the padded grid)."""
    return np.sqrt(N / P0) * np.pi / (4 * np.sqrt(2))


def ellipse_sinogram(ell, N0, P0, cor=None):
    """float64 [P0, N0] exact line integrals (unit-disk lengths) of the ellipse
    set in lpTransform's geometry: angle a pi / P0, detector column x at
    s = -1 + 2 (x + N/2 - int(cor)) / (N - 1) on the padded grid N, and
    s = x cos(th) - y sin(th) (x along columns, y along rows)."""
    import math
    cor = N0 / 2 if cor is None else cor
    N = int(math.ceil((N0 + abs(N0 / 2 - cor) * 2.0) / 16.0) * 16)
    th = np.arange(P0) * np.pi / P0
    s = -1 + 2 * (np.arange(N0) + N // 2 - int(cor)) / (N - 1)
    w = np.stack([np.cos(th), -np.sin(th)], axis=1)          # [P0, 2]
    out = np.zeros((P0, N0))
    for A, a, b, x0, y0, psi in ell:
        # in the ellipse's frame the line x.w = s meets the unit circle after
        # scaling by (a, b): chord = 2 a b sqrt(h^2 - d^2) / h^2
        wu = w[:, 0] * np.cos(psi) + w[:, 1] * np.sin(psi)
        wv = -w[:, 0] * np.sin(psi) + w[:, 1] * np.cos(psi)
        h2 = (a * wu) ** 2 + (b * wv) ** 2
        d = s[None, :] - (x0 * w[:, 0] + y0 * w[:, 1])[:, None]
        out += 2 * A * a * b * np.sqrt(np.clip(h2[:, None] - d * d, 0, None)) / h2[:, None]
    return out


def noisy_sinogram(p, photons, rng):
    """Transmission noise: counts ~ Poisson(photons e^-p), clamped to >= 1,
    returned as -log(counts / photons)."""
    counts = np.maximum(rng.poisson(photons * np.exp(-np.asarray(p, np.float64))), 1)
    return -np.log(counts / photons)


def scan(nslices, N0, P0, n_ellipses, seed, photons=None, cor=None):
    """-> (phantoms [ns, N0, N0] f32, sinograms [ns, P0, N0] f32 in fwdp units):
    seeded ellipse phantoms inside the unit disk, their exact sinograms with
    optional Poisson noise, scaled by radon_scale."""
    import math
    rng = np.random.default_rng(seed)
    c = N0 / 2 if cor is None else cor
    N = int(math.ceil((N0 + abs(N0 / 2 - c) * 2.0) / 16.0) * 16)
    imgs, sinos = [], []
    # the user's image is the centre N0 x N0 block of the padded grid, whose
    # pixels sit at linspace(-1, 1, N)
    lin = -1 + 2 * (np.arange(N0) + N // 2 - N0 // 2) / (N - 1)
    for _ in range(nslices):
        ell = random_ellipses(n_ellipses, rng, inside=True)
        imgs.append(rasterize(ell, N0, lin))
        p = ellipse_sinogram(ell, N0, P0, cor)
        if photons:
            p = noisy_sinogram(p, photons, rng)
        sinos.append(p * radon_scale(N, P0))
    return np.stack(imgs).astype(np.float32), np.stack(sinos).astype(np.float32)
