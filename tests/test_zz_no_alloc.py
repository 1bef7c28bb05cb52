"""lp_fwd_noalloc (the forward operator on a static per-GPU scratch array)
against lp_fwd through the C ABI: identical results on two calls in a row,
the second with fewer slices, and a clear error for calls whose scratch
exceeds the static array (rejected before anything is launched)."""
import ctypes

import numpy as np
import pytest

from lprec_amd import _lib, lpTransform
from lprec_amd.phantom import phantom_stack

pytestmark = pytest.mark.gpu


def _call(lib, name, h, out, inp, ns0, stream):
    return getattr(lib, name)(h, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(inp.data_ptr()), int(ns0),
                              ctypes.c_void_p(stream))


def test_fwd_noalloc_matches_fwd(torch_cuda):
    torch = torch_cuda
    lib = _lib.load()
    N0, P0, ns = 256, 384, 8
    lp = lpTransform(N0, P0, ns, "none", N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    h = lp.clphandle[0].h
    f = torch.from_numpy(phantom_stack(ns, N0, 12, seed=7)).cuda()
    stream = torch.cuda.current_stream().cuda_stream
    for ns0 in (8, 3):
        print("lp_fwd_noalloc vs lp_fwd, N0 %d, %d slices" % (N0, ns0), flush=True)
        R_ref = torch.zeros((ns0, P0, N0), device="cuda")
        R_na = torch.full((ns0, P0, N0), 7.0, device="cuda")
        fin = f[:ns0].contiguous()
        assert _call(lib, "lp_fwd", h, R_ref, fin, ns0, stream) == _lib.LP_OK
        assert _call(lib, "lp_fwd_noalloc", h, R_na, fin, ns0, stream) == _lib.LP_OK
        torch.cuda.synchronize()
        a, b = R_na.cpu().numpy(), R_ref.cpu().numpy()
        diff = float(np.max(np.abs(a - b)))
        print("max abs difference %.3e" % diff, flush=True)
        assert np.array_equal(a, b)


def test_fwd_noalloc_rejects_calls_beyond_the_static_scratch(torch_cuda):
    torch = torch_cuda
    lib = _lib.load()
    N0, P0 = 1024, 64
    lp = lpTransform(N0, P0, 2, "none", N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    h = lp.clphandle[0].h
    f = torch.zeros((2, N0, N0), device="cuda")
    R = torch.zeros((2, P0, N0), device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    # one slice at N = 1024 needs 52 N^2 = 54.5 MB (fits 64 MiB); two do not
    rc = _call(lib, "lp_fwd_noalloc", h, R, f, 2, stream)
    msg = lib.lp_last_error(h).decode()
    print("two slices at N = 1024: rc %d, %s" % (rc, msg), flush=True)
    assert rc == _lib.LP_ERR_ARG and "static scratch" in msg
    assert _call(lib, "lp_fwd_noalloc", h, R, f, 1, stream) == _lib.LP_OK
    torch.cuda.synchronize()
