"""tomopy-style entry point: reconstruct a stack of sinograms with one call.

The reference points tomopy users to a wrapper (README.rst:32-35) that builds
one ``lpTransform`` for a chunk of slices, initialises it on every GPU of
``gpu_list`` and runs an ``lpmethods`` function per chunk on a thread per GPU
(the pattern of tests/test_gpus_many.py:11-84 and
tests/test_gpus_many_map.py:14-100).  ``lprec`` below does the same with a
dynamic queue of slice chunks: each GPU thread takes the next chunk when it
finishes one, so uneven chunk costs balance out.
"""
import concurrent.futures as cf
import math
import queue

import numpy as np

from . import lpmethods
from .lpTransform import lpTransform


def check_uniform_theta(theta, nproj):
    """The log-polar method assumes nproj angles k*pi/nproj on [0, pi)
    (initsgl.py:36); anything else is rejected."""
    if theta is None:
        return
    theta = np.asarray(theta, dtype=np.float64).ravel()
    if theta.size != nproj:
        raise ValueError("theta has %d angles, the sinogram %d" % (theta.size, nproj))
    ref = np.arange(nproj) * np.pi / nproj
    if np.max(np.abs(theta - ref)) > 1e-3 * np.pi / nproj:
        raise ValueError("lprec needs uniformly spaced angles k*pi/Nproj on [0, pi)")


def lprec(tomo, center, recon=None, theta=None, lpmethod="fbp", filter_name="parzen", interp_type="cubic",
          num_iter=64, reg_par=-1, gpu_list=(0,), chunk=None, make_transform=None):
    """Reconstruct ``tomo`` [Nslices, Nproj, N] (float32, host) -> [Nslices, N, N].

    ``center`` is the rotation centre (a scalar or a per-slice array; the
    first entry is used, as one operator serves the whole stack).  ``lpmethod``
    names a function of ``lprec_amd.lpmethods``; ``num_iter`` / ``reg_par`` are
    passed to it.  ``chunk`` = slices per operator call (default: as many as
    fit 2^24 pixels, the reference's choice).  ``make_transform(Nslices)`` may
    supply the operator (tests use a host operator); by default it is an
    ``lpTransform`` initialised on every GPU of ``gpu_list``.
    """
    tomo = np.ascontiguousarray(tomo, dtype=np.float32)
    if tomo.ndim != 3:
        raise ValueError("tomo must be [Nslices, Nproj, N]")
    ns, nproj, n = tomo.shape
    check_uniform_theta(theta, nproj)
    if lpmethod not in lpmethods.METHODS:
        raise ValueError("unknown lpmethod %r (one of %s)" % (lpmethod, ", ".join(lpmethods.METHODS)))
    method = lpmethods.METHODS[lpmethod]
    gpu_list = list(gpu_list)
    c = float(np.asarray(center, dtype=np.float64).ravel()[0])
    if chunk is None:
        chunk = max(1, min(int(2 ** 24 / float(n * n)), int(math.ceil(ns / float(len(gpu_list))))))
    chunk = max(1, min(int(chunk), ns))
    if recon is None:
        recon = np.zeros((ns, n, n), dtype=np.float32)
        if lpmethod == "em":
            recon += 1e-3
    recon = np.ascontiguousarray(recon, dtype=np.float32)

    if make_transform is None:
        lp = lpTransform(n, nproj, chunk, filter_name, c, interp_type)
        lp.precompute(lpmethod != "fbp")
        for gpu in gpu_list:
            lp.initcmem(lpmethod != "fbp", gpu)
    else:
        lp = make_transform(chunk)

    work = queue.Queue()
    for s0 in range(0, ns, chunk):
        work.put((s0, min(ns, s0 + chunk)))

    def run(gpu):
        done = 0
        while True:
            try:
                s0, s1 = work.get_nowait()
            except queue.Empty:
                return done
            recon[s0:s1] = method(lp, recon[s0:s1], tomo[s0:s1], num_iter, reg_par, gpu)
            done += s1 - s0

    with cf.ThreadPoolExecutor(len(gpu_list)) as ex:
        counts = list(ex.map(run, gpu_list))
    if sum(counts) != ns:
        raise RuntimeError("slice chunks were lost")
    return recon
