"""Build oracle/libref32.so from ref32.c (test infrastructure).

gcc with -ffp-contract=off -fno-fast-math: every float operator rounds once,
fmaf only where written, matching the HIP kernels' evaluation order.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "ref32.c")
OUT = os.path.join(HERE, "libref32.so")


def build(verbose=True):
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math",
           "-o", OUT, SRC, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("ref32 build failed:\n" + r.stderr)
    if verbose:
        print(f"built {OUT}", file=sys.stderr)
    return OUT


if __name__ == "__main__":
    build()
