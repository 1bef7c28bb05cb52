"""The row and column passes that the library compiles at run time
(lprec_amd/csrc/rtc_kernels.hpp, hipRTC) compile for gfx950 from the source
tree, without a GPU: a header change that breaks them fails here rather than
silently leaving a GPU run on the generic kernels.  The instantiations are
the ones a handle asks for at N = 2064 (h = 1032 = 2^3 3 43, rho 4128) and
N = 528 (h = 264 = 2^3 3 11, rho 1056), with the default variant, and the
filtered sinogram pair pass at N = 2064, for each number of odd-radix output
pairs per work item the library may choose (4, 6, 8)."""
import ctypes
import ctypes.util
import os

import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lprec_amd", "csrc")


def _hiprtc():
    for name in ("/opt/rocm/lib/libhiprtc.so", ctypes.util.find_library("hiprtc")):
        if name:
            try:
                return ctypes.CDLL(name)
            except OSError:
                pass
    pytest.skip("libhiprtc not available")


NAMES = [
    "lpct::k_rows_fwd_ct<1032, 8, 1032, 1024, 4, 1, 43, 8, 3>",
    "lpct::k_rows_inv_ct<1032, 8, 1032, 1024, 4, 1, 43, 8, 3>",
    "lpct::k_cols_ct<4128, 512, 4, 2, 43, 8, 4, 3>",
    "lpct::k_rows_fwd_ct<264, 8, 264, 256, 4, 1, 11, 8, 3>",
    "lpct::k_cols_ct<1056, 256, 4, 2, 11, 8, 4, 3>",
    "lpct::k_sino_rows_pair_ct<2064, 512, 4, 43, 8, 4, 3>",
]


@pytest.mark.parametrize("pp", [4, 6, 8])
def test_runtime_passes_compile_for_gfx950(pp):
    rtc = _hiprtc()
    prog = ctypes.c_void_p()
    src = b'#include "rtc_kernels.hpp"\n'
    assert rtc.hiprtcCreateProgram(ctypes.byref(prog), src, b"lprec_rtc.hip", 0, None, None) == 0
    for n in NAMES:
        assert rtc.hiprtcAddNameExpression(prog, n.encode()) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", ("-I" + CSRC).encode(), b"-DLPCT_ODD_PP=%d" % pp,
            b"-Xclang", b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]
    arr = (ctypes.c_char_p * len(opts))(*opts)
    rc = rtc.hiprtcCompileProgram(prog, len(opts), arr)
    if rc != 0:
        n = ctypes.c_size_t()
        rtc.hiprtcGetProgramLogSize(prog, ctypes.byref(n))
        log = ctypes.create_string_buffer(n.value + 1)
        rtc.hiprtcGetProgramLog(prog, log)
        pytest.fail("hipRTC compile failed:\n" + log.value.decode(errors="replace")[:3000])
    for n in NAMES:
        low = ctypes.c_char_p()
        assert rtc.hiprtcGetLoweredName(prog, n.encode(), ctypes.byref(low)) == 0 and low.value
    size = ctypes.c_size_t()
    assert rtc.hiprtcGetCodeSize(prog, ctypes.byref(size)) == 0 and size.value > 0
    rtc.hiprtcDestroyProgram(ctypes.byref(prog))
