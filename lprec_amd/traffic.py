"""Algorithmic HBM bytes of the log-polar pipeline (the B_alg model), shared by
the benchmark's headline, multi-scan and iterative lines.

Compulsory traffic of a fused design per slice, FP32, cubic interpolation:
user input read once and output written once; the B-spline prefilter as two
sweeps of one read and one write of the padded array (the first sweep fused
with padding / angle stuffing, the stuffed zero rows neither written nor
read on the adjoint side); the gather source read once; per span the 2-D
log-polar convolution as 4 L (L = 4 Nrho Ntheta bytes: R2C writes the half
spectra, the column pass reads and writes them, C2R reads them); tables and
kernel spectra shared by a batch and not counted.
"""
import math


def geometry(N0, P0, cor=None):
    """Internal sizes of lpTransform(N0, P0, ..., cor): N (initsgl.py:29),
    os with Python's banker's rounding and P = os P0 (initsgl.py:32-33),
    Nrho = 2N, Ntheta = N."""
    cor = N0 / 2 if cor is None else cor
    N = int(math.ceil((N0 + abs(N0 / 2 - cor) * 2.0) / 16.0) * 16)
    os_ = max(int(round(3.0 * N / 2.0 / P0)), 1)
    return {"N0": N0, "Nproj0": P0, "N": N, "os": os_, "Nproj": os_ * P0, "Nrho": 2 * N, "Ntheta": N}


def b_fwd(g):
    """4 N0^2 + 16 N^2 + 12 L + 4 N0 P0."""
    N0, N, P0 = g["N0"], g["N"], g["Nproj0"]
    L = 4 * g["Nrho"] * g["Ntheta"]
    return 4 * N0 * N0 + 16 * N * N + 12 * L + 4 * N0 * P0


def b_adj(g, filtered=False):
    """4 N0 P0 + 8 N P0 + 8 N P + 12 L + 4 N0^2 (the cubic path writes only the
    kept angle rows of the stuffed sinogram); with an FBP filter the first
    sweep writes all P rows: 4 N0 P0 + 16 N P + 12 L + 4 N0^2."""
    N0, N, P0, P = g["N0"], g["N"], g["Nproj0"], g["Nproj"]
    L = 4 * g["Nrho"] * g["Ntheta"]
    if filtered:
        return 4 * N0 * P0 + 16 * N * P + 12 * L + 4 * N0 * N0
    return 4 * N0 * P0 + 8 * N * P0 + 8 * N * P + 12 * L + 4 * N0 * N0


def b_alg(g):
    """B_alg per slice of one fwd + adj pair."""
    return b_fwd(g) + b_adj(g)


def bytes_per_slice(N0, P0, cor=None):
    """B_alg per slice of fwd + adj for a scan geometry."""
    return b_alg(geometry(N0, P0, cor))
