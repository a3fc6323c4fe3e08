"""Build the CPU oracle (oracle/libce_oracle.so) with g++.

Test infrastructure only: loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg.
Compiled without floating-point contraction (fma only where the source calls it) so that its FP64
results are reproducible bit for bit by the device kernels.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "ce_oracle.cpp")
LIB = os.path.join(HERE, "libce_oracle.so")


def build(verbose=True):
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-mfma", "-mavx2", "-fno-math-errno", "-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
