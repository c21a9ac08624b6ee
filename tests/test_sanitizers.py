"""The CPU suite under AddressSanitizer and UndefinedBehaviorSanitizer.

Builds the oracle (oracle/Makefile target liboracle_san.so) and the host side
of the C ABI layer (api.hip's builders, path walking and ensemble copies:
bacon_amd.build.build_sanitized) with -fsanitize=address,undefined, then runs
every other CPU test in a child process that loads those builds
(BACON_ORACLE_LIB, BACON_GPU_LIB) with the clang sanitizer runtime preloaded.
Any sanitizer report aborts the child and fails this test.
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _asan_runtime():
    c = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return c[-1] if c else None


def test_cpu_suite_under_sanitizers():
    import oracle
    from bacon_amd import build as B
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("no clang AddressSanitizer runtime in this image")
    if not all(os.path.exists(os.path.join(B.OBJ, os.path.splitext(s)[0] + ".o"))
               for s in B.SOURCES if s != "api.hip"):
        pytest.skip("the regular build's objects are absent (run bacon_amd.build first)")
    san_oracle = oracle.build(force=True, target="liboracle_san.so")
    san_gpu = B.build_sanitized(verbose=False)
    env = dict(os.environ)
    env.update(LD_PRELOAD=rt, BACON_ORACLE_LIB=san_oracle, BACON_GPU_LIB=san_gpu,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    cmd = [sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-m", "not gpu", "-q",
           "-x", "-p", "no:cacheprovider",
           "--ignore", os.path.join(ROOT, "tests", "test_sanitizers.py")]
    try:
        r = subprocess.run(cmd, env=env, cwd=ROOT, capture_output=True, text=True, timeout=900)
    finally:
        # rebuilt on every run; not left in the tree that ships to GPU runs
        for p in (san_oracle, san_gpu):
            if os.path.exists(p):
                os.remove(p)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert " passed" in out
