"""Drop-in ``lpTransform`` (src/lprec/lpTransform.py:9-50) on MI355X.

Same constructor, method names, argument order and device-pointer I/O as the
reference.  Arrays are torch (ROCm) tensors, or any object exposing
``__cuda_array_interface__``; they must be C-contiguous float32 device arrays
on the handle's GPU.  Unlike the reference, shapes, dtypes, devices and
``Ns0 <= Nslices`` are validated (``ValueError``), filter names are matched
case-insensitively (the reference's own tests pass 'None', which its
initsadj.py:107 would reject), and failures raise instead of resetting the
device.
"""
import os

import numpy as np

from . import _lib, precache

FILTERS = ("none", "ramp", "shepp-logan", "cosine", "cosine2", "hamming", "hann", "parzen")


def normalize_filter(filter_type):
    f = "none" if filter_type is None else str(filter_type).lower()
    if f not in FILTERS:
        raise ValueError("unknown filter_type %r (one of %s)" % (filter_type, ", ".join(FILTERS)))
    return f


def padded_size(N0, cor):
    """Internal grid size (initsgl.py:29)."""
    import math
    return int(math.ceil((N0 + abs(N0 / 2 - cor) * 2.0) / 16.0) * 16)


def _device_array(a, shape, name, gpu):
    """-> (data pointer, stream) of a validated float32 C-contiguous device array."""
    shape = tuple(int(s) for s in shape)
    try:
        import torch
    except ImportError:  # pragma: no cover - torch is part of the stack
        torch = None
    if torch is not None and isinstance(a, torch.Tensor):
        if not a.is_cuda:
            raise ValueError("%s must be a GPU tensor" % name)
        if a.device.index != gpu:
            raise ValueError("%s is on cuda:%d, the handle is on cuda:%d" % (name, a.device.index, gpu))
        if a.dtype != torch.float32:
            raise ValueError("%s must be float32, got %s" % (name, a.dtype))
        if not a.is_contiguous():
            raise ValueError("%s must be C-contiguous" % name)
        if tuple(a.shape) != shape:
            raise ValueError("%s has shape %s, expected %s" % (name, tuple(a.shape), shape))
        return a.data_ptr(), torch.cuda.current_stream(gpu).cuda_stream
    cai = getattr(a, "__cuda_array_interface__", None)
    if cai is None:
        raise ValueError("%s must be a torch GPU tensor or expose __cuda_array_interface__" % name)
    if np.dtype(cai["typestr"]) != np.float32:
        raise ValueError("%s must be float32" % name)
    if tuple(cai["shape"]) != shape:
        raise ValueError("%s has shape %s, expected %s" % (name, tuple(cai["shape"]), shape))
    if cai.get("strides") not in (None, tuple(np.array([int(np.prod(shape[i + 1:])) * 4 for i in range(len(shape))]))):
        raise ValueError("%s must be C-contiguous" % name)
    stream = cai.get("stream", None)
    return int(cai["data"][0]), (0 if stream in (None, 1, 2) else int(stream))


class lpTransform:
    """Log-polar Radon transform ``fwdp`` and its adjoint ``adjp`` (drop-in)."""

    def __init__(self, N, Nproj, Nslices, filter_type, cor, interp_type):
        if int(N) != N or N < 16:
            raise ValueError("N must be an integer >= 16")
        if int(Nproj) != Nproj or Nproj < 3:
            raise ValueError("Nproj must be an integer >= 3")
        if int(Nslices) != Nslices or Nslices < 1:
            raise ValueError("Nslices must be an integer >= 1")
        self.N = int(N)
        self.Nslices = int(Nslices)
        self.filter_type = normalize_filter(filter_type)
        self.cor = cor
        self.Nproj = int(Nproj)
        self.interp_type = interp_type
        self.clphandle = [None] * 16
        self._fwd_ready = False

    @property
    def cubic(self):
        # initsgl.py:55 -- 'cubic' selects B-splines, anything else bilinear
        return self.interp_type == "cubic"

    def precompute(self, flg, cache_dir=None):
        """Build the operator tables on the host (lpTransform.py:19-31).

        With ``cache_dir`` (or the environment variable LPREC_AMD_CACHE) the
        tables are loaded from / stored to an on-disk cache keyed by the
        geometry (lprec_amd.precache)."""
        cache_dir = cache_dir or os.environ.get("LPREC_AMD_CACHE") or None
        pre = None
        if cache_dir:
            pre = precache.load(cache_dir, self.N, self.Nproj, self.Nslices, self.cor, self.cubic,
                                self.filter_type, bool(flg))
        if pre is None:
            pre = _lib.precompute(self.N, self.Nproj, self.Nslices, self.cor, self.cubic, self.filter_type,
                                  bool(flg))
            if cache_dir:
                precache.store(cache_dir, self.N, self.Nproj, self.cor, self.cubic, self.filter_type, pre)
        self.glpars = pre["glpars"]
        if flg:
            self.fwdparsi = pre["fwdparsi"]
            self.fwdparamsf = pre["fwdparamsf"]
            self._fwd_ready = True
        self.adjparsi = pre["adjparsi"]
        self.adjparamsf = pre["adjparamsf"]

    def initcmem(self, flg, gpu, schedule=None):
        """Create the per-GPU handle and upload the tables (lpTransform.py:34-40).

        ``schedule``: optional {knob: value} applied before the tables are
        uploaded (see lprec_amd.tune.KNOBS; "boxcap" and "boxstride" only take
        effect here)."""
        if not 0 <= int(gpu) < len(self.clphandle):
            raise ValueError("gpu index must be in [0, 16)")
        if not hasattr(self, "glpars"):
            raise ValueError("call precompute() before initcmem()")
        if flg and not self._fwd_ready:
            raise ValueError("precompute(1) is required before initcmem(1, gpu)")
        h = _lib.Handle(self.glpars, gpu)
        if schedule:
            from .tune import KNOBS
            for k, v in schedule.items():
                h.set_tuning(KNOBS[k], int(v))
        if flg:
            h.init_fwd(self.fwdparsi, self.fwdparamsf)
        h.init_adj(self.adjparsi, self.adjparamsf)
        old = self.clphandle[gpu]
        self.clphandle[gpu] = h
        if old is not None:
            old.close()

    def _handle(self, gpu):
        if not 0 <= int(gpu) < len(self.clphandle) or self.clphandle[gpu] is None:
            raise ValueError("no handle on gpu %r: call initcmem(flg, gpu) first" % (gpu,))
        return self.clphandle[gpu]

    @property
    def Nproj0(self):
        return self.Nproj

    def _ns0(self, a, name):
        shape = a.shape if hasattr(a, "shape") else a.__cuda_array_interface__["shape"]
        if len(shape) != 3:
            raise ValueError("%s must be 3-D [Nslices0, ., .]" % name)
        ns0 = int(shape[0])
        if not 1 <= ns0 <= self.Nslices:
            raise ValueError("%s holds %d slices; the transform was built for 1..%d" % (name, ns0, self.Nslices))
        return ns0

    def fwdp(self, R, f, gpu):
        """R[Ns0, Nproj, N] = Radon(f[Ns0, N, N]) (lpTransform.py:42-45)."""
        h = self._handle(gpu)
        ns0 = self._ns0(f, "f")
        fp, stream = _device_array(f, (ns0, self.N, self.N), "f", gpu)
        Rp, _ = _device_array(R, (ns0, self.Nproj, self.N), "R", gpu)
        h.fwd(Rp, fp, ns0, stream)

    def adjp(self, f, R, gpu):
        """f[Ns0, N, N] = back-projection of R[Ns0, Nproj, N], FBP-filtered if a
        filter was chosen (lpTransform.py:47-50)."""
        h = self._handle(gpu)
        ns0 = self._ns0(R, "R")
        Rp, stream = _device_array(R, (ns0, self.Nproj, self.N), "R", gpu)
        fp, _ = _device_array(f, (ns0, self.N, self.N), "f", gpu)
        h.adj(fp, Rp, ns0, stream)

    def fwdp_transpose(self, f, R, gpu):
        """f[Ns0, N, N] = F^T R[Ns0, Nproj, N]: the exact transpose of fwdp
        (no reference counterpart; the gradient of fwdp)."""
        h = self._handle(gpu)
        ns0 = self._ns0(R, "R")
        Rp, stream = _device_array(R, (ns0, self.Nproj, self.N), "R", gpu)
        fp, _ = _device_array(f, (ns0, self.N, self.N), "f", gpu)
        h.fwd_T(fp, Rp, ns0, stream)

    def adjp_transpose(self, R, f, gpu):
        """R[Ns0, Nproj, N] = A^T f[Ns0, N, N]: the exact transpose of adjp
        (FBP filter included; no reference counterpart)."""
        h = self._handle(gpu)
        ns0 = self._ns0(f, "f")
        fp, stream = _device_array(f, (ns0, self.N, self.N), "f", gpu)
        Rp, _ = _device_array(R, (ns0, self.Nproj, self.N), "R", gpu)
        h.adj_T(Rp, fp, ns0, stream)

    def close(self):
        for i, h in enumerate(self.clphandle):
            if h is not None:
                h.close()
                self.clphandle[i] = None
