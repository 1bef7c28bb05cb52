"""CPU oracle for the log-polar Radon transform path.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``lprec_amd`` package imports this
package: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker (never as the thing that
is measured or shipped).

Contents
--------
``precompute``  FP64 numpy restatement of the reference's host precompute
                (src/lprec/initsgl.py, initsfwd.py, initsadj.py) producing the
                reference's flat parameter arrays (the wire format read by
                src/cuda/params.cu:40-145).
``mirror``      FP64 numpy restatement of the reference's device pipeline
                (src/cuda/lpRgpu.cu:184-382 with the kernels of
                src/include/main_kernels.cuh and simple_kernels.cuh), fed by
                those same flat arrays, with exact bicubic B-spline weights.
``radon``       FP64 analytic ellipse-phantom Radon transform and a direct
                (pixel-driven) back-projection, used as physics references.

Pins (what each function is checked against) are listed in each module's
header; a function that no independent check covers says "parity unpinned".
"""
