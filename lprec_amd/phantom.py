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


def rasterize(ell, N):
    """float64 [N, N] image of the ellipse set, masked to the unit disk."""
    lin = np.linspace(-1, 1, N)
    X, Y = np.meshgrid(lin, lin)
    img = np.zeros((N, N))
    for A, a, b, x0, y0, psi in ell:
        c, s = np.cos(psi), np.sin(psi)
        u = (X - x0) * c + (Y - y0) * s
        v = -(X - x0) * s + (Y - y0) * c
        img[(u / a) ** 2 + (v / b) ** 2 <= 1] += A
    img[X ** 2 + Y ** 2 > 1] = 0
    return img


def phantom_stack(nslices, N, n_ellipses, seed, inside=False):
    """float32 [nslices, N, N]: one independent seeded phantom per slice."""
    rng = np.random.default_rng(seed)
    return np.stack([rasterize(random_ellipses(n_ellipses, rng, inside), N)
                     for _ in range(nslices)]).astype(np.float32)
