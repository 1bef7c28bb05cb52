"""On-disk cache of the precomputed operator tables.

The reference rebuilds every table for each ``lpTransform`` with Python loops
(initsfwd.py:85-87, initsadj.py:194-199); here the five flat arrays of the
wire format (initsfwd.py:96-126, initsadj.py:205-244) are stored once per
geometry as an ``.npz`` file and reloaded.  The key is (N0, Nproj, cor,
interp, filter) plus a format version; ``Nslices`` is not part of the key (it
only sits in ``glpars[6]``, which is patched on load).  Files are written
atomically and read with ``allow_pickle=False``.
"""
import os

import numpy as np

FORMAT = 1


def cache_path(cache_dir, N0, Nproj, cor, cubic, filter_type):
    cor_key = float(cor).hex().replace(".", "_").replace("+", "p").replace("-", "m")
    name = "lp_v%d_N%d_P%d_c%s_%s_%s.npz" % (FORMAT, int(N0), int(Nproj), cor_key,
                                            "cubic" if cubic else "linear", filter_type)
    return os.path.join(cache_dir, name)


def load(cache_dir, N0, Nproj, Nslices, cor, cubic, filter_type, want_fwd):
    """-> dict of the five arrays, or None when the cache has no entry that
    covers the request (a fwd request needs the fwd tables)."""
    path = cache_path(cache_dir, N0, Nproj, cor, cubic, filter_type)
    if not os.path.exists(path):
        return None
    try:
        with np.load(path, allow_pickle=False) as z:
            if int(z["format"]) != FORMAT:
                return None
            if want_fwd and "fwdparsi" not in z.files:
                return None
            out = {k: np.array(z[k]) for k in ("glpars", "adjparsi", "adjparamsf")}
            out["fwdparsi"] = np.array(z["fwdparsi"]) if "fwdparsi" in z.files else None
            out["fwdparamsf"] = np.array(z["fwdparamsf"]) if "fwdparamsf" in z.files else None
    except (OSError, ValueError, KeyError):
        return None
    out["glpars"][6] = np.float32(Nslices)
    if not want_fwd:
        out["fwdparsi"] = out["fwdparamsf"] = None
    return out


def store(cache_dir, N0, Nproj, cor, cubic, filter_type, pre):
    os.makedirs(cache_dir, exist_ok=True)
    path = cache_path(cache_dir, N0, Nproj, cor, cubic, filter_type)
    arrays = {k: v for k, v in pre.items() if v is not None}
    arrays["format"] = np.array(FORMAT)
    tmp = "%s.tmp%d" % (path, os.getpid())
    with open(tmp, "wb") as fh:
        np.savez(fh, **arrays)
    os.replace(tmp, path)
    return path
