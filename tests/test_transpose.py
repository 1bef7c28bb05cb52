"""Pins of the oracle's exact transposes (oracle/transpose.py), CPU only.

Each stage and both whole operators pass the dot-product test
<A x, y> = <x, A^T y> in FP64 against the mirror's forward operators, and the
prefilter transpose equals the explicit transposed matrix of the reference
recursion (main_kernels.cuh:97-142), for lines shorter and longer than its
12-term causal horizon.
"""
import numpy as np
import pytest

from oracle import mirror, precompute as pc
from oracle import transpose as tr


def _dot_rel(a, b):
    return abs(a - b) / max(abs(a), abs(b))


@pytest.mark.parametrize("L", [7, 12, 40])
def test_prefilter_transpose_is_matrix_transpose(L):
    M = mirror.prefilter_line(np.eye(L), 0)          # column j = P e_j
    MT = tr.prefilter_line_T(np.eye(L), 0)
    assert np.allclose(MT, M.T, atol=1e-14, rtol=0)


@pytest.mark.parametrize("cubic", [True, False])
def test_interp_transpose_dot(cubic):
    rng = np.random.default_rng(0)
    H, W = 13, 17
    T = rng.standard_normal((2, H, W))
    x = rng.uniform(-3, W + 3, 300)
    y = rng.uniform(-3, H + 3, 300)
    v = rng.standard_normal((2, 300))
    lhs = np.sum(mirror.interp(T, x, y, cubic) * v)
    rhs = np.sum(T * tr.interp_T(v, x, y, H, W, cubic))
    assert _dot_rel(lhs, rhs) < 1e-13


def test_convolution_transpose_dot():
    rng = np.random.default_rng(1)
    F = rng.standard_normal((2, 12, 8))
    G = rng.standard_normal((2, 12, 8))
    fZ = rng.standard_normal((12, 5)) + 1j * rng.standard_normal((12, 5))
    lhs = np.sum(mirror.lp_convolve(F, fZ) * G)
    rhs = np.sum(F * tr.lp_convolve_T(G, fZ))
    assert _dot_rel(lhs, rhs) < 1e-13


@pytest.mark.parametrize("filt,cor", [("none", 32), ("shepp-logan", 29.5), ("parzen", 35)])
def test_sinogram_preparation_transpose_dot(filt, cor):
    pre = pc.precompute(64, 40, 1, cor, "cubic", filt, fwd=False)
    tab = mirror.Tables(pre["glpars"], None, None, pre["adjparsi"], pre["adjparamsf"])
    rng = np.random.default_rng(2)
    R = rng.standard_normal((2, 40, 64))
    Y = rng.standard_normal((2, tab.Nproj, tab.N))
    lhs = np.sum(mirror.sino_prepare(tab, R) * Y)
    rhs = np.sum(R * tr.sino_prepare_T(tab, Y, 40))
    assert _dot_rel(lhs, rhs) < 1e-12


GEOMS = [(64, 48, 32, "cubic", "none"), (64, 40, 30.5, "linear", "none"), (48, 64, 25, "cubic", "hamming")]


@pytest.mark.parametrize("N0,P0,cor,interp,filt", GEOMS)
def test_operator_transposes_dot(N0, P0, cor, interp, filt):
    pre = pc.precompute(N0, P0, 1, cor, interp, filt)
    tab = mirror.Tables(pre["glpars"], pre["fwdparsi"], pre["fwdparamsf"], pre["adjparsi"], pre["adjparamsf"])
    rng = np.random.default_rng(3)
    f = rng.standard_normal((2, N0, N0))
    R = rng.standard_normal((2, P0, N0))
    # fwd_T is the transpose of fwd
    lhs = np.sum(mirror.fwd(tab, f) * R)
    rhs = np.sum(f * tr.fwd_T(tab, R))
    assert _dot_rel(lhs, rhs) < 1e-12
    # adj_T is the transpose of adj
    lhs = np.sum(mirror.adj(tab, R) * f)
    rhs = np.sum(R * tr.adj_T(tab, f, P0))
    assert _dot_rel(lhs, rhs) < 1e-12
    # the reference pair itself is only approximately adjoint
    faithful = _dot_rel(np.sum(mirror.fwd(tab, f) * R), np.sum(f * mirror.adj(tab, R)))
    assert faithful > 1e-6


def test_transposes_are_linear_and_slice_independent():
    pre = pc.precompute(48, 36, 1, 24, "cubic", "none")
    tab = mirror.Tables(pre["glpars"], pre["fwdparsi"], pre["fwdparamsf"], pre["adjparsi"], pre["adjparamsf"])
    rng = np.random.default_rng(4)
    R = rng.standard_normal((2, 36, 48))
    a = tr.fwd_T(tab, R)
    b = tr.fwd_T(tab, R[1:])
    assert np.allclose(a[1:], b, atol=1e-12)
    assert np.allclose(tr.fwd_T(tab, 3 * R), 3 * a, atol=1e-11)
    assert np.array_equal(tr.fwd_T(tab, 0 * R), np.zeros_like(a))
