"""The kernels' integer quotients by runtime divisors (FDiv in
lprec_amd/csrc/fft_lds.hpp) come from a float reciprocal: q = trunc((t + 0.5)
* fl(1/d)) with IEEE round-to-nearest float32 operations.  Checked here
exhaustively against integer division for every divisor up to 4200 and every
t below 2^15 (the index ranges of the FFT stages and row passes), and on
random pairs up to the stated bound t < 2^22 - 1."""
import numpy as np


def _fdiv(t, d):
    tf = t.astype(np.float32) + np.float32(0.5)
    inv = np.float32(1.0) / np.float32(d)
    return np.trunc(tf * inv).astype(np.int64)


def test_fdiv_exhaustive_small():
    t = np.arange(1 << 15, dtype=np.int64)
    for d in range(1, 4201):
        q = _fdiv(t, d)
        assert np.array_equal(q, t // d), d


def test_fdiv_random_up_to_bound():
    rng = np.random.default_rng(5)
    t = rng.integers(0, (1 << 22) - 1, size=1 << 20)
    for d in list(rng.integers(1, 1 << 24, size=64)) + [3, 43, 129, 1032, 2064, 4128, 8256, (1 << 22) - 3]:
        q = _fdiv(t, int(d))
        assert np.array_equal(q, t // int(d)), d
