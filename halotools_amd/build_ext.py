"""Build libhalopair.so in-tree with hipcc for gfx950 (MI355X).

The library is loaded through ctypes (halotools_amd/_engine.py); it is plain
HIP C++ with an extern "C" ABI (include/halopair.h) and no PyTorch types.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib", "libhalopair.so")
SOURCES = ["cellbuild.hip", "pairs.hip", "reduce.hip", "capi.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # bit-exact FP64: the reference engines evaluate dx*dx + dy*dy + dz*dz
    # without fused multiply-add, so no contraction anywhere
    "-ffp-contract=off",
    # keep hipcc from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False, extra=None, out=None):
    out = out or OUT
    os.makedirs(os.path.dirname(out), exist_ok=True)
    cmd = [HIPCC] + FLAGS + (extra or []) + [os.path.join(SRC, s) for s in SOURCES] + ["-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
