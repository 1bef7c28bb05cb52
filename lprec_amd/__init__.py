"""lprec_amd: log-polar Radon transform and back-projection on AMD MI355X.

``lpTransform`` is the drop-in counterpart of the reference's
src/lprec/lpTransform.py; the compute path is liblprec_amd.so (hand-written
HIP for gfx950) reached through ctypes.
"""
from .lpTransform import lpTransform, normalize_filter, padded_size  # noqa: F401
from ._lib import LprecError  # noqa: F401

__version__ = "0.1.0"
