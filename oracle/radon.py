"""FP64 physics references (TEST INFRASTRUCTURE).

``ellipse_radon``  analytic Radon transform of ellipse phantoms in the
                   geometry of the reference's fwd: detector s = x cos(th) -
                   y sin(th) with x along columns and y along rows of
                   linspace(-1, 1, N), th_a = a pi / P0 (established from
                   initsfwd.py:20-56: the LP->Cartesian map with the y sign
                   flip, and the polar grid proj = a pi/P - beta/2 measured
                   from the span rotations k beta + beta/2).
``lprec_radon_scale``  the factor between the reference fwd output and line
                   integrals in unit-disk length: the fwd kernel constant
                   sqrt(N os / P) pi / 4 / aR / sqrt(2) (initsfwd.py:57-58)
                   times aR (the log-polar frame is the unit disk scaled by
                   aR, initsfwd.py:24-27), i.e. sqrt(N / P0) pi / (4 sqrt 2).
``direct_backprojection``  pixel-driven back-projection sum_a R(th_a, x.w_a) dth.
"""
import numpy as np


def ellipse_radon(ell, N, P0, s=None):
    """[P0, N] line integrals of the ellipse set (A, a, b, x0, y0, psi)."""
    th = np.arange(P0) * np.pi / P0
    s = np.linspace(-1, 1, N) if s is None else s
    out = np.zeros((P0, len(s)))
    wx, wy = np.cos(th), -np.sin(th)          # s = x cos th - y sin th
    for A, a, b, x0, y0, psi in ell:
        eu = (np.cos(psi), np.sin(psi))
        ev = (-np.sin(psi), np.cos(psi))
        h2 = a * a * (wx * eu[0] + wy * eu[1]) ** 2 + b * b * (wx * ev[0] + wy * ev[1]) ** 2
        c = x0 * wx + y0 * wy
        d = s[None, :] - c[:, None]
        inside = d * d <= h2[:, None]
        out += np.where(inside, 2 * A * a * b * np.sqrt(np.maximum(h2[:, None] - d * d, 0)) / h2[:, None], 0)
    return out


def lprec_radon_scale(N, P0):
    return np.sqrt(N / P0) * np.pi / (4 * np.sqrt(2))


def direct_backprojection(R, N):
    """sum over angles of R(th_a, x cos th_a - y sin th_a) * pi / P0 (linear in s)."""
    P0 = R.shape[0]
    lin = np.linspace(-1, 1, N)
    X, Y = np.meshgrid(lin, lin)
    th = np.arange(P0) * np.pi / P0
    out = np.zeros((N, N))
    for a in range(P0):
        sp = X * np.cos(th[a]) - Y * np.sin(th[a])
        out += np.interp(sp.ravel(), lin, R[a], left=0, right=0).reshape(N, N)
    return out * np.pi / P0
