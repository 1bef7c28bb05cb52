"""Pins of the CPU oracle (oracle/precompute.py, oracle/mirror.py).

Each check is independent of the formula it checks: scipy routines, closed
forms, brute-force sums, the reference's own fixtures and tolerances
(tests/run_test.py:14-26 of the reference), and an analytic Radon transform.
A planted mistake in the oracle (sign, index, dropped term) fails at least
one of them.
"""
import os

import numpy as np
import pytest
from scipy import ndimage

from oracle import mirror, precompute as pc
from oracle.radon import ellipse_radon, lprec_radon_scale
from lprec_amd.phantom import random_ellipses, rasterize

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_prefilter_matches_scipy_spline_filter():
    # main_kernels.cuh:97-142 == B-spline prefilter with half-sample mirror
    # boundaries, up to the 12-term causal horizon (|z|^13 ~ 4e-8)
    rng = np.random.default_rng(0)
    a = rng.standard_normal((7, 64))
    b = mirror.prefilter_line(a, -1)
    c = ndimage.spline_filter1d(a, order=3, axis=-1, mode="reflect")
    assert np.max(np.abs(b - c)) / np.max(np.abs(c)) < 2e-7
    # interior is exact
    assert np.max(np.abs(b[:, 20:-20] - c[:, 20:-20])) < 1e-12


def test_prefilter_then_interp_reproduces_interior_samples():
    rng = np.random.default_rng(1)
    T = rng.standard_normal((40, 50))
    C = mirror.prefilter2d(T)
    yy, xx = np.mgrid[5:35, 5:45]
    v = mirror.interp(C, xx.ravel().astype(float), yy.ravel().astype(float), True)
    assert np.max(np.abs(v - T[5:35, 5:45].ravel())) < 1e-6


def test_interp_matches_scipy_map_coordinates_grid_wrap():
    # both forms of the sampler: the numpy restatement and the C one
    # (oracle/interp.c), which the mirror runs when it is built
    assert mirror._clib() is not None, "oracle/liboracle.so not built (python -m oracle.build)"
    rng = np.random.default_rng(2)
    T = rng.standard_normal((2, 37, 53))
    x = rng.uniform(-60, 120, 2000)
    y = rng.uniform(-60, 100, 2000)
    for cubic, order in ((True, 3), (False, 1)):
        for fn in (mirror.interp_taps, mirror.interp):
            a = fn(T, x, y, cubic)
            for s in range(2):
                b = ndimage.map_coordinates(T[s], [y, x], order=order, mode="grid-wrap", prefilter=False)
                assert np.max(np.abs(a[s] - b)) < 1e-12


def test_interp_brute_force_16_taps():
    rng = np.random.default_rng(3)
    T = rng.standard_normal((9, 11))

    def b3(t):  # centred cubic B-spline
        t = abs(t)
        return (2 / 3 - t * t + t ** 3 / 2) if t < 1 else ((2 - t) ** 3 / 6 if t < 2 else 0.0)

    for x, y in ((4.3, 2.8), (-0.6, 8.7), (10.9, -3.2)):
        ix, iy = int(np.floor(x)), int(np.floor(y))
        fx, fy = x - ix, y - iy
        ref = sum(b3(fx - (a - 1)) * b3(fy - (b - 1)) * T[(iy - 1 + b) % 9, (ix - 1 + a) % 11]
                  for a in range(4) for b in range(4))
        for fn in (mirror.interp_taps, mirror.interp):
            assert abs(fn(T, [x], [y])[0] - ref) < 1e-12


def test_lp_convolve_is_circular_convolution():
    rng = np.random.default_rng(4)
    F = rng.standard_normal((8, 6))
    K = rng.standard_normal((8, 6))
    fZ = np.fft.rfft2(K)
    ref = np.zeros_like(F)
    for i in range(8):
        for j in range(6):
            ref[i, j] = sum(F[a, b] * K[(i - a) % 8, (j - b) % 6] for a in range(8) for b in range(6))
    assert np.allclose(mirror.lp_convolve(F, fZ), ref, atol=1e-12)


def test_apply_filter_is_centred_convolution():
    # applyFilter == circular convolution of the 4N-padded row with the kernel
    # whose spectrum is ifftshift(filter)
    rng = np.random.default_rng(5)
    N = 16
    filt = rng.uniform(0, 1, 4 * N)
    row = rng.standard_normal((1, 1, N))
    out = mirror.apply_filter(row, filt)
    k = np.real(np.fft.ifft(np.fft.ifftshift(filt)))
    M = 4 * N
    ref = np.array([sum(row[0, 0, j] * k[(i - j) % M] for j in range(N)) for i in range(N)])
    assert np.allclose(out[0, 0], ref, atol=1e-12)


def test_b3com_closed_form():
    P, _ = pc.create_gl(64, 96, 1, 32, "cubic")
    for grid, M in ((P.thsp, P.Ntheta), (P.rhosp, P.Nrho)):
        B = np.fft.fft(np.fft.ifftshift(pc.splineB3(grid, 1)))
        k = np.arange(M)
        # the float32 lattice is uniform only to float32 rounding, which moves
        # the tap distances off 1 and 2 by ~1e-6 of a step
        assert np.max(np.abs(B - (4 + 2 * np.cos(2 * np.pi * k / M)) / 6)) < 1e-5


def test_fzeta_rows_match_direct_quadrature():
    # fzeta row kappa, frequency nu = sum_m h_m cos(t_m)^e exp(-2 pi i nu m / Nl) * dt
    Nt, Nr = 32, 64
    betas, rhos = 2 * np.pi / 3, 1.317
    for adjoint in (False, True):
        F = pc.fzeta(Nt, Nr, betas, rhos, 0, 4, adjoint, rows=[5, 40])
        Nl = 4 * Nt
        t = np.arange(-Nl / 2, Nl / 2) / Nl * betas
        h = pc._h_weights(Nl)
        for r, j in enumerate([5, 40]):
            kap = j - Nr / 2
            e = (2j * np.pi * kap / rhos) if adjoint else (-2j * np.pi * kap / rhos - 1)
            g = np.fft.fftshift(h * np.cos(t) ** e)     # sample m <-> theta = m dt (periodic)
            for c in (1, 7, Nt - 3):
                nu = c - Nt / 2
                m = np.arange(Nl)
                ref = np.sum(g * np.exp(-2j * np.pi * nu * m / Nl)) * (t[1] - t[0])
                assert abs(F[r, c] - ref) < 1e-10 * max(1, abs(ref))


def test_wint_core_is_exact_for_polynomials():
    # sum w_k g(t_k) == int t g(t) dt over [t0, t_end] for deg(g) < 12
    t = np.arange(0, 257) / 512.0
    w = pc.wint_core(12, t)
    a, b = t[0], t[-1]
    # (exact up to the ~1e-9 conditioning of the 12x12 Vandermonde inverse)
    for d in (0, 1, 2, 5, 8, 11):
        exact = (b ** (d + 2) - a ** (d + 2)) / (d + 2)
        assert abs(np.sum(w * t ** d) - exact) < 2e-9 * exact, d
    # a non-polynomial integrand is only approximated
    exact = (np.sin(30 * b) / 900 - b * np.cos(30 * b) / 30) - 0.0
    assert 1e-12 < abs(np.sum(w * np.sin(30 * t)) - exact) < 1e-6


def test_wint_tail_is_linear_ramp():
    t = np.arange(0, 257) / 512.0
    w, wc = pc.wint(12, t), pc.wint_core(12, t)
    N = len(t)
    assert np.array_equal(w[:-40], wc[:-40])
    assert np.allclose(w[-40:], wc[-40] / (N - 40) * np.arange(N - 40, N), rtol=0, atol=1e-18)
    # interior weights approximate t dt (the ramp filter |nu|)
    dt = t[1] - t[0]
    assert np.allclose(w[20:-45], t[20:-45] * dt, rtol=1e-6)


def test_geometry_invariants():
    pre = pc.precompute(256, 384, 1, 128, "cubic", "none")
    fw, aj, P = pre["fwd"], pre["adj"], pre["geom"]
    # spans partition the sinogram (initsfwd.py:38-40)
    allp = np.concatenate(fw["pids"])
    assert len(allp) == P.N * P.Nproj and len(np.unique(allp)) == len(allp)
    # Appendix-A counts for config 1 (fwd) and the disk membership
    assert len(fw["cids"]) == 42917
    assert len(aj["cids"]) == 51040
    assert len(np.intersect1d(aj["lpids"], aj["wids"])) == 0
    # span membership uses the float32 theta lattice: every span starts at theta = -beta/2
    th = P.thsp[aj["lpids"] % P.Ntheta]
    assert th.min() == np.float32(-P.beta / 2)


def test_filter_shapes_and_symmetry():
    for name in pc.FILTERS:
        f = pc.take_filter(64, name)
        assert f.shape == (256,)
        # centred real even response: f[128 + m] == f[128 - m]
        assert np.allclose(f[129:], f[127:0:-1])
        assert np.all(f >= 0)
    with pytest.raises(ValueError):
        pc.normalize_filter("gauss")
    assert pc.normalize_filter("None") == "none"


def _fixture(ns):
    N, P0 = 512, 384
    f = np.fromfile(os.path.join(GOLDEN, "f"), dtype=np.float32).reshape(1, N, N)
    R = np.fromfile(os.path.join(GOLDEN, "R"), dtype=np.float32).reshape(1, P0, N)
    fa = np.concatenate([f * (ns - k) for k in range(ns)]).astype(np.float32)
    Ra = np.concatenate([R * (ns - k) for k in range(ns)]).astype(np.float32)
    return fa, Ra


@pytest.fixture(scope="module")
def fixture_tables():
    pre = pc.precompute(512, 384, 2, 256, "cubic", "None")
    return mirror.Tables(pre["glpars"], pre["fwdparsi"], pre["fwdparamsf"], pre["adjparsi"], pre["adjparamsf"])


def test_reference_fixture_test_adj_many(fixture_tables):
    # test_adj_many.py:9-65 with the thresholds of run_test.py:18-20
    fa, Ra = _fixture(2)
    tab = fixture_tables
    Rf = mirror.fwd(tab, fa)
    fR = mirror.adj(tab, Ra)
    RfR = mirror.fwd(tab, fR.astype(np.float32))
    err0 = abs(np.sum(Rf * Ra) - np.sum(fR * fa)) / abs(np.sum(fR * fa))
    err1 = np.linalg.norm(fR - fa) / np.linalg.norm(fa)
    err2 = np.linalg.norm(RfR - Rf) / np.linalg.norm(Rf)
    assert err0 < 0.005 and err1 < 0.007 and err2 < 0.0007, (err0, err1, err2)


def test_reference_norm_adj_fwd(fixture_tables):
    # test_lpmethods.py:47-52 + run_test.py:25: norm of adj(fwd(f)) over 8
    # identical slices of the fixture f is 222.4238 (reference, one-sided +10)
    fa, _ = _fixture(1)
    tab = fixture_tables
    back = mirror.adj(tab, mirror.fwd(tab, fa))
    norm8 = np.sqrt(8) * np.linalg.norm(back)
    assert abs(norm8 - 222.4238) < 2.0, norm8


def test_fwd_matches_analytic_ellipse_radon():
    # physics pin: disk-contained ellipses, analytic line integrals
    rng = np.random.default_rng(11)
    N, P0 = 256, 384
    ell = random_ellipses(6, rng, inside=True)
    f = rasterize(ell, N)[None]
    pre = pc.precompute(N, P0, 1, N / 2, "cubic", "none")
    tab = mirror.Tables(pre["glpars"], pre["fwdparsi"], pre["fwdparamsf"], pre["adjparsi"], pre["adjparamsf"])
    R = mirror.fwd(tab, f)[0]
    Ra = ellipse_radon(ell, N, P0) * lprec_radon_scale(N, P0)
    # pixelised phantom vs exact ellipses: discretisation error of the raster
    assert np.linalg.norm(R - Ra) / np.linalg.norm(Ra) < 0.03


def test_product_sinogram_generator_matches_oracle_radon():
    # lprec_amd.phantom.ellipse_sinogram (synthetic bench/test data) against
    # the oracle's analytic Radon transform, centred and with a centre offset
    from lprec_amd.phantom import ellipse_sinogram
    rng = np.random.default_rng(12)
    ell = random_ellipses(7, rng, inside=True)
    assert np.allclose(ellipse_sinogram(ell, 128, 90), ellipse_radon(ell, 128, 90), atol=1e-12)
    # cor = 61 -> padded N = 144, detector column x at s = -1 + 2 (x + 72 - 61) / 143
    s = -1 + 2 * (np.arange(128) + 72 - 61) / 143
    assert np.allclose(ellipse_sinogram(ell, 128, 90, cor=61), ellipse_radon(ell, 128, 90, s=s), atol=1e-12)


def _bubbles():
    # tests/data/fbub, Rbub of the reference (make_bubles.m:1-19): three discs
    # at N = 256, 64 angles; Rbub holds the exactly Hamming-filtered Radon data
    # written column-major as [angle][s] (so [s][angle] in C order), fbub the
    # exactly filtered image
    N, P0 = 256, 64
    R = np.fromfile(os.path.join(GOLDEN, "Rbub"), dtype=np.float32).reshape(N, P0).T.copy()
    f = np.fromfile(os.path.join(GOLDEN, "fbub"), dtype=np.float32).reshape(N, N)
    return N, P0, R, f


def _fitted_error(x, ref):
    a = float(np.sum(x * ref) / np.sum(x * x))
    return float(np.linalg.norm(a * x - ref) / np.linalg.norm(ref))


def test_reference_bubbles_fixture_backprojection():
    # the unfiltered back-projection of the pre-filtered data against the
    # filtered image, up to the fixture's own scale: no reference test states
    # a tolerance; measured 0.217 (64 angles undersample N = 256).  The image
    # axis convention is pinned too: the data only match the image flipped
    # along its rows, every other orientation is far off.
    N, P0, R, f = _bubbles()
    pre = pc.precompute(N, P0, 1, N / 2, "cubic", "none", fwd=False)
    tab = mirror.Tables(pre["glpars"], None, None, pre["adjparsi"], pre["adjparamsf"])
    x = mirror.adj(tab, R[None].astype(np.float64))[0]
    err = _fitted_error(x, f[::-1])
    others = [_fitted_error(x, g) for g in (f, f.T, f[:, ::-1], f[::-1, ::-1])]
    assert err < 0.25, err
    assert min(others) > 0.5, others
