"""Compiles the oracle's C restatement (petch_oracle.c) with gcc into oracle/_build/.
Test infrastructure only."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libpetch_oracle.so")
SRC = os.path.join(HERE, "petch_oracle.c")


def build(verbose=False):
    if not os.path.exists(SRC):
        return None
    os.makedirs(OUT, exist_ok=True)
    # -ffp-contract=off: products are rounded one at a time, as the kernels round them
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-Wall", "-Wno-unused-function", SRC, "-o", LIB + ".tmp", "-lm"]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    if verbose and p.stdout:
        print(p.stdout)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
