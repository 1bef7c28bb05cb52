"""ctypes binding of liblprec_amd.so (the C ABI in include/lprec_amd.h).

The library is built in-tree by ``python -m lprec_amd.build``.  Importing this
module does not need a GPU; every device entry point reports a HIP error
through its return code when no device is present.  There is no CPU fallback:
when the library is missing, ``load()`` raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LPREC_AMD_LIB overrides the in-tree library (e.g. a build with other flags)
LIB_PATH = os.environ.get("LPREC_AMD_LIB") or os.path.join(HERE, "liblprec_amd.so")

LP_OK = 0
LP_ERR_ARG = -1
LP_ERR_ALLOC = -2
LP_ERR_DEVICE = -3
LP_ERR_UNSUPPORTED = -4

# every symbol declared in include/lprec_amd.h
EXPORTS = (
    "lp_pre_create", "lp_pre_sizes", "lp_pre_copy", "lp_pre_destroy", "lp_pre_error",
    "lp_create", "lp_init_fwd", "lp_init_adj", "lp_fwd", "lp_adj", "lp_fwd_T", "lp_adj_T", "lp_destroy",
    "lp_last_error", "lp_mem_info", "lp_info", "lp_set_tuning", "lp_get_tuning",
    "lp_profile", "lp_profile_read", "lp_debug_stop", "lp_debug_copy", "lp_stream_bench", "lp_fwd_noalloc",
)


class LprecError(RuntimeError):
    """A failed call into liblprec_amd (carries the library's message)."""

    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


_lib = None


def load():
    """Load (once) and return the ctypes library; raises ImportError if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("liblprec_amd.so not built: run `python -m lprec_amd.build` (%s)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp, i32, i64p = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)
    sig = {
        "lp_pre_create": (i32, [i32, i32, i32, ctypes.c_double, i32, ctypes.c_char_p, i32, ctypes.POINTER(vp)]),
        "lp_pre_sizes": (i32, [vp, i64p]),
        "lp_pre_copy": (i32, [vp, vp, vp, vp, vp, vp]),
        "lp_pre_destroy": (None, [vp]),
        "lp_pre_error": (ctypes.c_char_p, []),
        "lp_create": (i32, [vp, i32, ctypes.POINTER(vp)]),
        "lp_init_fwd": (i32, [vp, vp, vp]),
        "lp_init_adj": (i32, [vp, vp, vp]),
        "lp_fwd": (i32, [vp, vp, vp, i32, vp]),
        "lp_fwd_noalloc": (i32, [vp, vp, vp, i32, vp]),
        "lp_adj": (i32, [vp, vp, vp, i32, vp]),
        "lp_fwd_T": (i32, [vp, vp, vp, i32, vp]),
        "lp_adj_T": (i32, [vp, vp, vp, i32, vp]),
        "lp_destroy": (None, [vp]),
        "lp_last_error": (ctypes.c_char_p, [vp]),
        "lp_mem_info": (i32, [i32, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
        "lp_info": (i32, [vp, i64p]),
        "lp_set_tuning": (i32, [vp, i32, i32]),
        "lp_get_tuning": (i32, [vp, i32, ctypes.POINTER(ctypes.c_int)]),
        "lp_profile": (i32, [vp, i32]),
        "lp_profile_read": (i32, [vp, i32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)]),
        "lp_debug_stop": (i32, [vp, i32]),
        "lp_debug_copy": (i32, [vp, i32, vp, ctypes.c_size_t]),
        "lp_stream_bench": (i32, [i32, vp, vp, ctypes.c_size_t, i32, i32, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def precompute(N0, Nproj, Nslices, cor, cubic, filter_type, want_fwd):
    """Host precompute (C++) -> dict of the five wire arrays (numpy)."""
    lib = load()
    h = ctypes.c_void_p()
    filt = (filter_type or "none").encode()
    rc = lib.lp_pre_create(int(N0), int(Nproj), int(Nslices), float(cor), int(bool(cubic)), filt,
                           int(bool(want_fwd)), ctypes.byref(h))
    if rc != LP_OK:
        msg = lib.lp_pre_error().decode()
        if rc == LP_ERR_ARG:
            raise ValueError(msg)
        raise LprecError(rc, msg)
    try:
        sz = (ctypes.c_int64 * 5)()
        lib.lp_pre_sizes(h, sz)
        out = {
            "glpars": np.zeros(sz[0], np.float32),
            "fwdparsi": np.zeros(sz[1], np.int32) if want_fwd else None,
            "fwdparamsf": np.zeros(sz[2], np.float32) if want_fwd else None,
            "adjparsi": np.zeros(sz[3], np.int32),
            "adjparamsf": np.zeros(sz[4], np.float32),
        }
        lib.lp_pre_copy(h, _ptr(out["glpars"]), _ptr(out["fwdparsi"]), _ptr(out["fwdparamsf"]),
                        _ptr(out["adjparsi"]), _ptr(out["adjparamsf"]))
    finally:
        lib.lp_pre_destroy(h)
    return out


def stream_bench(device, dst_ptr, src_ptr, n_floats, kind, blocks, stream):
    """HBM streaming kernel (0 copy, 1 read, 2 write) on raw device pointers."""
    rc = load().lp_stream_bench(int(device), ctypes.c_void_p(dst_ptr), ctypes.c_void_p(src_ptr), int(n_floats),
                                int(kind), int(blocks), ctypes.c_void_p(stream))
    if rc != LP_OK:
        raise LprecError(rc, "lp_stream_bench failed")


class Handle:
    """Owns one lp_handle (device tables + scratch for one geometry)."""

    def __init__(self, glpars, device):
        self.lib = load()
        self.h = ctypes.c_void_p()
        glpars = np.ascontiguousarray(glpars, dtype=np.float32)
        rc = self.lib.lp_create(_ptr(glpars), int(device), ctypes.byref(self.h))
        if rc != LP_OK:
            msg = self.lib.lp_last_error(None).decode()
            self.h = None
            raise LprecError(rc, msg)
        self.device = int(device)

    def _check(self, rc):
        if rc != LP_OK:
            raise LprecError(rc, self.lib.lp_last_error(self.h).decode())

    def init_fwd(self, parsi, parsf):
        parsi = np.ascontiguousarray(parsi, dtype=np.int32)
        parsf = np.ascontiguousarray(parsf, dtype=np.float32)
        self._check(self.lib.lp_init_fwd(self.h, _ptr(parsi), _ptr(parsf)))

    def init_adj(self, parsi, parsf):
        parsi = np.ascontiguousarray(parsi, dtype=np.int32)
        parsf = np.ascontiguousarray(parsf, dtype=np.float32)
        self._check(self.lib.lp_init_adj(self.h, _ptr(parsi), _ptr(parsf)))

    def fwd(self, R_ptr, f_ptr, ns0, stream):
        self._check(self.lib.lp_fwd(self.h, ctypes.c_void_p(R_ptr), ctypes.c_void_p(f_ptr), int(ns0),
                                    ctypes.c_void_p(stream)))

    def adj(self, f_ptr, R_ptr, ns0, stream):
        self._check(self.lib.lp_adj(self.h, ctypes.c_void_p(f_ptr), ctypes.c_void_p(R_ptr), int(ns0),
                                    ctypes.c_void_p(stream)))

    def fwd_T(self, f_ptr, R_ptr, ns0, stream):
        self._check(self.lib.lp_fwd_T(self.h, ctypes.c_void_p(f_ptr), ctypes.c_void_p(R_ptr), int(ns0),
                                      ctypes.c_void_p(stream)))

    def adj_T(self, R_ptr, f_ptr, ns0, stream):
        self._check(self.lib.lp_adj_T(self.h, ctypes.c_void_p(R_ptr), ctypes.c_void_p(f_ptr), int(ns0),
                                      ctypes.c_void_p(stream)))

    def info(self):
        v = (ctypes.c_int64 * 12)()
        self._check(self.lib.lp_info(self.h, v))
        keys = ("N", "N0", "Ntheta", "Nrho", "Nproj", "Nproj0", "os", "Nslices", "cor", "cubic",
                "has_filter", "device_bytes")
        return dict(zip(keys, list(v)))

    def set_tuning(self, key, value):
        self._check(self.lib.lp_set_tuning(self.h, int(key), int(value)))

    def get_tuning(self, key):
        v = ctypes.c_int()
        self._check(self.lib.lp_get_tuning(self.h, int(key), ctypes.byref(v)))
        return v.value

    STAGES = {0: ("prep_rows", "col_iir", "gather_lp", "rows_fwd", "cols", "rows_inv", "polar_out"),
              1: ("sino_rows", "col_iir", "gather_lp", "rows_fwd", "cols", "rows_inv", "cart_out")}

    def profile(self, enable):
        """Enable (and clear) or disable per-stage HIP-event timing."""
        self._check(self.lib.lp_profile(self.h, int(bool(enable))))

    def profile_read(self, which):
        """-> (dict stage -> summed ms, number of calls) for which = 0 fwd / 1 adj."""
        ms = (ctypes.c_float * 7)()
        n = ctypes.c_int()
        self._check(self.lib.lp_profile_read(self.h, int(which), ms, ctypes.byref(n)))
        return dict(zip(self.STAGES[which], list(ms))), n.value

    def debug_stop(self, stage):
        """Return from the next fwd/adj calls after `stage` (0 = full run)."""
        self._check(self.lib.lp_debug_stop(self.h, int(stage)))

    def debug_copy(self, which, shape, dtype=np.float32):
        """Copy internal buffer `which` (0 fc, 1 Rc, 2 grids, 3 spectra) into a new array."""
        out = np.empty(shape, dtype=dtype)
        self._check(self.lib.lp_debug_copy(self.h, int(which), out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
        return out

    def close(self):
        if self.h:
            self.lib.lp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
