"""Builds the oracle (test infrastructure: the CPU parity checker) into oracle/liboracle.so with gcc.

IEEE arithmetic, no fast-math, no floating-point contraction (fused multiply-adds are explicit fmaf calls).

    python oracle/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(quiet=True):
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-Wall",
           "-o", tmp, os.path.join(HERE, "oracle.c"), os.path.join(HERE, "hploc.c"), "-lm", "-lpthread"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    if r.stdout.strip() and not quiet:
        print(r.stdout)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    try:
        print(build(quiet=False))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
