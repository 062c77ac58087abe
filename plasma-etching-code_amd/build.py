"""Build recipe for libpetch_hip.so (gfx950) and the CPU oracle library.

    python plasma-etching-code_amd/build.py            # both libraries
    python plasma-etching-code_amd/build.py --hip-only

hipcc cross-compiles for gfx950 without a GPU. Every call rebuilds everything (objects go to
plasma-etching-code_amd/_build/, the shared library into the petch package directory so the
ctypes loader finds it next to its Python wrappers).
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "petch", "libpetch_hip.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PETCH_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-munsafe-fp-atomics",
    # packed FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def hip_sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build_hip(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = hip_sources()

    def one(src):
        obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        if verbose:
            print("  compiled", os.path.basename(src), flush=True)
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(one, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose=False):
    """Compiles the C restatement under oracle/ (test infrastructure, never linked into the
    product library)."""
    sys.path.insert(0, ROOT)
    from oracle import build as ob   # noqa: E402
    return ob.build(verbose=verbose)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--hip-only", action="store_true")
    ap.add_argument("--oracle-only", action="store_true")
    ap.add_argument("-j", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    if not a.oracle_only:
        print("built", build_hip(a.j, a.v))
    if not a.hip_only:
        print("built", build_oracle(a.v))


if __name__ == "__main__":
    main()
