"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

Two sanitized builds: libhalopair with its host code instrumented (request
planning, bin tables, the C ABI entry points' validation;
``halotools_amd.build_ext.build_host_sanitized``) and the oracle's exact C
restatement (``make -C oracle san``), both with ROCm's clang so that one
sanitizer runtime serves them.  The host-logic and oracle test modules run
against them in a child process started with that runtime preloaded; any
sanitizer report fails the child (halt_on_error) and therefore this test.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _stale(out, sources):
    return not os.path.exists(out) or any(os.path.getmtime(s) > os.path.getmtime(out) for s in sources)


def test_host_and_oracle_suites_under_asan_ubsan():
    from halotools_amd import build_ext
    runtime = build_ext.sanitizer_runtime()
    if runtime is None or not os.path.exists(build_ext.HIPCC):
        pytest.skip("ROCm clang sanitizer runtime not found")
    srcs = [os.path.join(build_ext.SRC, s) for s in build_ext.SOURCES]
    srcs += [os.path.join(build_ext.SRC, h) for h in ("hp_internal.h", "pairs_kernel.h")]
    srcs.append(os.path.join(ROOT, "include", "halopair.h"))
    if _stale(build_ext.SAN_OUT, srcs):
        build_ext.build_host_sanitized()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=runtime, HALOPAIR_LIB=build_ext.SAN_OUT,
               HP_ORACLE_LIB=os.path.join(ROOT, "oracle", "_build", "libhp_oracle_san.so"),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
           os.path.join(ROOT, "tests", "test_host.py"), os.path.join(ROOT, "tests", "test_oracle.py")]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert " passed" in out
