"""Pins of the oracle's FBP filters (oracle/precompute.py take_filter, wint)
and of the fzeta end-correction weights, independent of their formulas.

* The reference's fixtures tests/data/R and Rbub are Hamming-filtered
  analytic sinograms written by MATLAB (make_phantom.m, make_bubles.m,
  apply_filter_exact.m).  Re-running that generator (tests/matlab_restate.py)
  with the oracle's take_filter(N, 'hamming') as the filter reproduces both
  files to float32 rounding; any other window misses by > 1e-2.  This pins
  the Hamming window, the wint quadrature weights of the ramp and the
  centring of the filter array.
* The fixture image tests/data/f is the same phantom filtered in 2-D
  (apply_filter_2d_exact.m); its restatement reproduces the file.  By the
  Fourier slice theorem the FBP with window W (the oracle's adj with filter
  W applied to the unfiltered analytic sinogram) equals the phantom filtered
  in 2-D by the radial window W: checked for every window, each closer to its
  own 2-D filtered phantom than to any other window's (hann and cosine2 are
  the same function for d = 1/2).
* Window values at t = 1/4 and t = 1/2 (the last tabulated frequency) equal
  their closed forms: hamming 0.54 / 0.08, hann 1/2 / 0, cosine sqrt(2)/2 / 0,
  cosine2 1/2 / 0, shepp-logan sinc(1/4) = 2 sqrt(2)/pi / 2/pi, parzen 1/8 / 0.
* The ramp FBP of analytic ellipse data, in the fixture's sinogram scale
  (make_phantom.m:14), reconstructs the rasterised phantom at absolute scale.
* fzeta's end-corrected trapezoid weights (the 11 integers over 958003200,
  initsfwd.py:78-84) integrate smooth even functions over [-beta, beta) --
  among them fzeta's own integrands cos(t)^(-+2 pi i kappa / rhos - ...) --
  to 1e-13 against scipy.integrate.quad, converge at high order, and
  changing any one integer by one unit breaks the 1e-12 level.
"""
import os

import numpy as np
import pytest

import matlab_restate as mr
from oracle import mirror, precompute as pc
from oracle.radon import ellipse_radon
from lprec_amd.phantom import random_ellipses, rasterize

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def adj_oracle(N, P0, filt, R):
    pre = pc.precompute(N, P0, 1, N / 2, "cubic", filt, fwd=False)
    tab = mirror.Tables(pre["glpars"], None, None, pre["adjparsi"], pre["adjparamsf"])
    return mirror.adj(tab, R[None])[0]


def test_hamming_filter_reproduces_the_fixture_sinograms():
    N, Nt = 512, 384
    gold = np.fromfile(os.path.join(GOLDEN, "R"), dtype=np.float32).reshape(Nt, N).astype(np.float64)
    scale = mr.fixture_sinogram_scale(N, Nt)
    errs = {}
    for name in ("hamming", "hann", "cosine", "shepp-logan", "ramp", "parzen"):
        h = mr.filtered_sinogram(N, Nt, mr.SHEPP_LOGAN, pc.take_filter(N, name))
        errs[name] = rel(h[:, ::-1] * scale, gold)
    print("sinogram fixture R vs regenerated with each window:", errs)
    assert errs["hamming"] < 1e-6
    assert min(v for k, v in errs.items() if k != "hamming") > 1e-2
    # bubbles fixture (N = 256, 64 angles, three discs; stored unscaled, [N][Ntheta])
    N, Nt = 256, 64
    gold = np.fromfile(os.path.join(GOLDEN, "Rbub"), dtype=np.float32).reshape(N, Nt).T.astype(np.float64)
    e = rel(mr.filtered_sinogram(N, Nt, mr.BUBBLES, pc.take_filter(N, "hamming")), gold)
    e_hann = rel(mr.filtered_sinogram(N, Nt, mr.BUBBLES, pc.take_filter(N, "hann")), gold)
    print("bubbles fixture Rbub: hamming %.2e, hann %.2e" % (e, e_hann))
    assert e < 1e-6 and e_hann > 1e-2


@pytest.fixture(scope="module")
def shepp_logan_512():
    N, Nt = 512, 384
    F = mr.image_spectrum(N, mr.SHEPP_LOGAN)
    h = mr.filtered_sinogram(N, Nt, mr.SHEPP_LOGAN, np.ones(4 * N))
    R = h[:, ::-1] * mr.fixture_sinogram_scale(N, Nt)
    return N, Nt, F, R


def test_2d_filter_restatement_reproduces_the_fixture_image(shepp_logan_512):
    N, _, F, _ = shepp_logan_512
    gold = np.fromfile(os.path.join(GOLDEN, "f"), dtype=np.float32).reshape(N, N).astype(np.float64)
    e = rel(mr.filtered_image(N, F, "hamming")[:, ::-1], gold)
    print("image fixture f vs regenerated: %.2e" % e)
    assert e < 1e-6


WINDOWS = ("hamming", "hann", "cosine", "cosine2", "shepp-logan", "parzen")
# FBP(window) vs the 2-D filtered phantom: errors measured 0.0052, 0.0036,
# 0.0078, 0.0036, 0.031, 0.0026 (sampling of the projections; ramp-like
# windows with a high cut-off keep the most aliasing)
BOUND = {"hamming": 0.008, "hann": 0.006, "cosine": 0.012, "cosine2": 0.006, "shepp-logan": 0.045, "parzen": 0.004}
SAME = {frozenset(("hann", "cosine2"))}


def test_fbp_windows_equal_2d_filtered_phantoms(shepp_logan_512):
    N, Nt, F, R = shepp_logan_512
    refs = {w: mr.filtered_image(N, F, w)[:, ::-1] for w in WINDOWS}
    for w in WINDOWS:
        x = adj_oracle(N, Nt, w, R)
        errs = {v: rel(x, refs[v]) for v in WINDOWS}
        print("FBP %-12s vs 2-D filtered phantoms: %s" % (w, ", ".join("%s %.4f" % kv for kv in errs.items())))
        assert errs[w] < BOUND[w]
        for v in WINDOWS:
            if v != w and frozenset((v, w)) not in SAME:
                assert errs[w] < errs[v], (w, v)


@pytest.mark.parametrize("N", [256, 512])
def test_window_values_at_closed_form_points(N):
    ramp = pc.take_filter(N, "ramp").astype(np.float64)
    c = 2 * N                      # centre of the length-4N array
    quarter, half = c + N, 0       # t = N / 4N = 1/4; index 0 holds t = 2N / 4N = 1/2
    expect = {"hamming": (0.54, 0.08), "hann": (0.5, 0.0), "cosine": (np.sqrt(0.5), 0.0), "cosine2": (0.5, 0.0),
              "shepp-logan": (2 * np.sqrt(2) / np.pi, 2 / np.pi), "parzen": (0.125, 0.0)}
    for name, (q, h) in expect.items():
        f = pc.take_filter(N, name).astype(np.float64)
        assert abs(f[quarter] / ramp[quarter] - q) < 1e-6, name
        assert abs(f[half] / ramp[half] - h) < 1e-6, name
        assert abs(f[c - N] - f[quarter]) == 0.0          # even about the centre


def test_ramp_fbp_absolute_scale():
    # the fixture's sinogram scale 2N sqrt(2) sqrt(N) / sqrt(Ntheta) (make_phantom.m:14)
    # turns line integrals into the data the reference's FBP inverts
    for N, P0 in ((256, 384), (512, 768)):
        ell = random_ellipses(8, np.random.default_rng(3), inside=True)
        f = rasterize(ell, N)
        R = ellipse_radon(ell, N, P0) * 2 * N * np.sqrt(2) * np.sqrt(N) / np.sqrt(P0)
        x = adj_oracle(N, P0, "ramp", R)
        X, Y = np.meshgrid(np.linspace(-1, 1, N), np.linspace(-1, 1, N))
        m = X ** 2 + Y ** 2 < 0.95 ** 2
        a = np.sum(x[m] * f[m]) / np.sum(x[m] * x[m])
        e = rel(x[m], f[m])
        print("ramp FBP N=%d: absolute error %.4f, best scale %.4f" % (N, e, 1 / a))
        assert abs(1 / a - 1) < 0.01 and e < 0.1


def _quad_even(g, beta):
    import warnings
    from scipy.integrate import quad
    with warnings.catch_warnings():
        # quad reports that it reached the rounding floor (~1e-16) before 1e-15
        warnings.simplefilter("ignore")
        re = quad(lambda t: g(t).real, -beta, beta, epsabs=1e-15, epsrel=1e-14, limit=400)[0]
        im = quad(lambda t: g(t).imag, -beta, beta, epsabs=1e-15, epsrel=1e-14, limit=400)[0]
    return re + 1j * im


def _rule(h, g, beta):
    Nl = len(h)
    t = np.arange(-Nl / 2, Nl / 2) / Nl * 2 * beta
    return np.sum(h * g(t)) * (t[1] - t[0])


def test_fzeta_end_correction_weights_match_quadrature():
    beta = np.pi / 3
    rhos = 1.317
    fns = [lambda t: np.exp(np.cos(3 * t)) * np.cosh(t),
           lambda t: 1.0 / (1.5 - np.cos(t))]
    for kap in (0.0, 1.0, 3.0):
        fns.append(lambda t, kap=kap: np.cos(t) ** (-2j * np.pi * kap / rhos - 1))   # fwd exponent, a = 0
        fns.append(lambda t, kap=kap: np.cos(t) ** (2j * np.pi * kap / rhos))        # adj exponent
    worst = 0.0
    for Nl in (512, 1024):
        h = pc._h_weights(Nl)
        for g in fns:
            ref = _quad_even(g, beta)
            worst = max(worst, abs(_rule(h, g, beta) - ref) / abs(ref))
    print("end-corrected trapezoid vs quad: worst relative error %.2e" % worst)
    assert worst < 1e-13
    # high order: doubling the points cuts the error of an oscillatory
    # integrand (kappa = 3) by more than 2^9
    g = fns[-2]
    ref = _quad_even(g, beta)
    e128, e256 = (abs(_rule(pc._h_weights(n), g, beta) - ref) / abs(ref) for n in (128, 256))
    print("kappa = 3: error %.2e at 128 points, %.2e at 256" % (e128, e256))
    assert e128 / e256 > 2 ** 9
    # one unit off in any of the 11 integers is visible
    ints = np.array([-216254335, 679543284, -1412947389, 2415881496, -3103579086, 2939942400, -2023224114,
                     984515304, -321455811, 63253516, -5675265], dtype=np.float64)
    Nl = 512
    g = fns[2]
    ref = _quad_even(g, beta)
    for j in range(len(ints)):
        bad = ints.copy()
        bad[j] += 1.0
        corr = 1 + bad / 958003200.0
        corr[0] = 2 * (corr[0] - 0.5)
        h = np.ones(Nl)
        h[0] *= corr[0]
        for i in range(1, len(corr)):
            h[i] *= corr[i]
            h[-i] *= corr[i]
        assert abs(_rule(h, g, beta) - ref) / abs(ref) > 1e-12, j
