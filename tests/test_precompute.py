"""Native (C++) precompute against the oracle restatement, on the CPU.

Integer index lists must be identical and float tables equal to float32
rounding; the global parameter array is bit-identical.  Error paths must raise
instead of crashing.
"""
import numpy as np
import pytest

from oracle import precompute as pc
from lprec_amd import _lib, lpTransform

GEOMS = [
    # N0, Nproj, cor, filter
    (256, 384, 128, "none"),
    (512, 384, 256, "hamming"),
    (528, 792, 261, "parzen"),      # N = 528 (factor 11), os = 1
    (250, 300, 127.3, "ramp"),      # centre offset: N = 256
    (200, 151, 100, "shepp-logan"),  # odd Nproj, os = round(1.99) = 2
    (96, 48, 48, "cosine2"),        # os = 3
    (2048, 1500, 1024, "none"),     # the headline workload's tables (N = 2048, os = 2)
    (2048, 1500, 1021.6, "ramp"),   # the FBP workload's centre offset: N = 2064 = 2^4 3 43
]


@pytest.mark.parametrize("N0,P0,cor,filt", GEOMS)
def test_native_precompute_matches_oracle(N0, P0, cor, filt):
    a = _lib.precompute(N0, P0, 2, cor, True, filt, True)
    o = pc.precompute(N0, P0, 2, cor, "cubic", filt)
    assert np.array_equal(a["glpars"], o["glpars"])
    for name in ("fwdparsi", "adjparsi"):
        assert np.array_equal(a[name], o[name]), name
    for name in ("fwdparamsf", "adjparamsf"):
        x, y = a[name].astype(np.float64), o[name].astype(np.float64)
        assert x.shape == y.shape
        scale = np.maximum(np.abs(y), 1e-3 * np.max(np.abs(y)))
        assert np.max(np.abs(x - y) / scale) < 3e-7, name


def test_native_precompute_linear_and_no_fwd():
    a = _lib.precompute(128, 192, 1, 64, False, "none", False)
    o = pc.precompute(128, 192, 1, 64, "linear", "none", fwd=False)
    assert a["fwdparsi"] is None
    assert np.array_equal(a["glpars"], o["glpars"])
    assert np.array_equal(a["adjparsi"], o["adjparsi"])


def test_osangles_banker_rounding():
    # initsgl.py:32 uses Python round(): 1.5 -> 2, 2.5 -> 2
    for N0, P0 in ((64, 64), (96, 57), (160, 96)):
        g = _lib.precompute(N0, P0, 1, N0 / 2, True, "none", False)["glpars"]
        assert int(g[8]) == pc.os_angles(int(g[0]), P0) == max(round(3.0 * int(g[0]) / 2.0 / P0), 1)


def test_precompute_errors():
    with pytest.raises(ValueError):
        _lib.precompute(256, 384, 1, 128, True, "gauss", True)
    with pytest.raises(ValueError):
        _lib.precompute(8, 384, 1, 4, True, "none", True)
    with pytest.raises(ValueError):
        lpTransform(256, 384, 1, "nope", 128, "cubic")
    lp = lpTransform(256, 384, 1, "None", 128, "cubic")   # 'None' is accepted (reference tests pass it)
    assert lp.filter_type == "none"
    with pytest.raises(ValueError):
        lp.initcmem(1, 0)   # precompute() first
