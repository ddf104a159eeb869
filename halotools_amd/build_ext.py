"""Build libhalopair.so in-tree with hipcc for gfx950 (MI355X).

The library is loaded through ctypes (halotools_amd/_engine.py); it is plain
HIP C++ with an extern "C" ABI (include/halopair.h) and no PyTorch types.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib", "libhalopair.so")
# the pair kernel's variants are instantiated one mode per translation unit
# (pairs_mode_*.hip) so that they compile in parallel
SOURCES = ["cellbuild.hip", "pairs.hip", "reduce.hip", "capi.hip"] + [
    "pairs_mode_%s.hip" % m for m in ("count3d", "proj", "xyz", "obj", "wxyz", "jack", "jackxy", "wsum", "jac", "tan")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # bit-exact FP64: the reference engines evaluate dx*dx + dy*dy + dz*dz
    # without fused multiply-add, so no contraction anywhere
    "-ffp-contract=off",
    # keep hipcc from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False, extra=None, out=None, jobs=None):
    """Compile every source to an object (in parallel), then link the library."""
    out = out or OUT
    objdir = out + ".obj"
    os.makedirs(objdir, exist_ok=True)
    flags = FLAGS + (extra or [])

    def compile_one(src):
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [HIPCC] + flags + ["-c", os.path.join(SRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    jobs = jobs or min(len(SOURCES), max(1, min(os.cpu_count() or 1, 16)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        objs = list(pool.map(compile_one, SOURCES))
    # roctx: named phase ranges for rocprofv3 --marker-trace
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-lrocprofiler-sdk-roctx", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


SAN_OUT = os.path.join(HERE, "_lib", "san", "libhalopair_san.so")
# host code only: the device code is built as usual
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]


def build_host_sanitized(out=None):
    """The library with its host code (request planning, bin tables, the C ABI)
    under AddressSanitizer and UndefinedBehaviorSanitizer, for CPU tests.

    Load it into a process started with ROCm clang's ASan runtime preloaded
    (``sanitizer_runtime()``) and ``HALOPAIR_LIB`` pointing at it.
    """
    return build(extra=SAN_FLAGS, out=out or SAN_OUT)


def sanitizer_runtime():
    """Path of the ASan runtime of ROCm's clang (shared; includes the UBSan handlers)."""
    import glob
    found = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return found[-1] if found else None


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
