"""The B_alg model (lprec_amd.traffic) on the CPU: its geometry equals the
native precompute's (glpars N, os, P), and its byte counts at the measured
workloads are the documented ones."""
import pytest

from lprec_amd import _lib, traffic


@pytest.mark.parametrize("N0,P0,cor", [(2048, 1500, 1024), (2048, 1500, 1021.6), (1024, 1024, 512), (512, 360, 256),
                                       (96, 57, 48), (160, 96, 80)])
def test_geometry_matches_native_precompute(N0, P0, cor):
    g = traffic.geometry(N0, P0, cor)
    gl = _lib.precompute(N0, P0, 1, cor, True, "none", False)["glpars"]
    assert (g["N"], g["os"], g["Nproj"], g["Ntheta"], g["Nrho"]) == (int(gl[0]), int(gl[8]), int(gl[5]), int(gl[2]),
                                                                    int(gl[3]))


def test_headline_bytes():
    g = traffic.geometry(2048, 1500)
    L = 4 * 4096 * 2048
    assert traffic.b_fwd(g) == 4 * 2048 ** 2 + 16 * 2048 ** 2 + 12 * L + 4 * 2048 * 1500
    assert traffic.b_alg(g) == 1004273664
    assert traffic.bytes_per_slice(2048, 1500) == traffic.b_alg(g)
    # FBP workload (centre offset, N = 2064): the filtered first sweep writes all P rows
    g2 = traffic.geometry(2048, 1500, 1021.6)
    assert g2["N"] == 2064
    assert traffic.b_adj(g2, filtered=True) > traffic.b_adj(g2)
