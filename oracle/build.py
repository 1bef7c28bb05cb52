"""Build recipe of the oracle's C restatement (TEST INFRASTRUCTURE).

``python -m oracle.build`` compiles ``oracle/interp.c`` with gcc into
``oracle/liboracle.so`` (plain IEEE double, no FMA contraction, OpenMP over
sample points).  Only the tests, ``smoke()`` and bench's CPU-baseline leg load
it, through ``oracle.mirror``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    tmp = OUT + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           os.path.join(HERE, "interp.c"), "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
