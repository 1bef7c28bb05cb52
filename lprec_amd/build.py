"""In-tree build of liblprec_amd.so (HIP kernels for gfx950 + host library).

``python -m lprec_amd.build`` or ``lprec_amd.build.build()``.  hipcc
cross-compiles for gfx950 without a GPU; the shared library is written next to
this file so that it travels with the source tree.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "liblprec_amd.so")
ARCH = os.environ.get("LPREC_AMD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # no packed-FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(verbose=False):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for src in ("kernels.hip", "handle.hip"):
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        _run([hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    # host precompute: plain IEEE double (no FMA contraction)
    obj = os.path.join(bdir, "precompute.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", "-c",
          os.path.join(CSRC, "precompute.cpp"), "-o", obj], verbose)
    objs.append(obj)
    tmp = OUT + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp, "-lpthread", "-lhiprtc", "-ldl"],
         verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
