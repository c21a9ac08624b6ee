"""Build libbacon_gpu.so in-tree: hipcc for gfx950, one object per source.

    python -m bacon_amd.build            # compile everything
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbacon_gpu.so")
ARCH = os.environ.get("BACON_GPU_ARCH", "gfx950")

SOURCES = ["bdf_kernels.hip", "bdf_systems.hip", "api.hip", "rk_kernels.hip",
           "adjoint_kernels.hip", "linear_kernels.hip", "tune.hip", "graph.hip",
           "adams_kernels.hip"]

# -ffp-contract=off: every multiply-add in the kernels is an explicit fma(),
# so results are reproducible operation by operation.
# -target-feature -packed-fp32-ops: no packed FP32 instructions are emitted.
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


CPP_TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "test_reference_ivp.cpp")
CPP_TEST_BIN = os.path.join(os.path.dirname(HERE), "tests", "cpp", "test_reference_ivp.bin")


def build_cpp_tests(verbose: bool = True) -> str:
    """The C++ restatement of the reference's ivp tests, linked against the
    library (run on a GPU by tests/test_gpu_cpp.py)."""
    cmd = ["g++", "-std=c++17", "-O2", "-o", CPP_TEST_BIN, CPP_TEST_SRC,
           f"-L{HERE}", "-lbacon_gpu", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../bacon_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {CPP_TEST_BIN}")
    return CPP_TEST_BIN


SAN_LIB = os.path.join(HERE, "libbacon_gpu_san.so")
# Host-side AddressSanitizer + UndefinedBehaviorSanitizer for the C ABI layer
# (builders, path walking, ensemble copies in api.hip); device code is built
# as usual.  Each -fsanitize= applies to the host pass only.
SAN_HOST_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                  "-Xarch_host", "-fno-sanitize-recover=undefined",
                  "-Xarch_host", "-fno-omit-frame-pointer", "-g"]


def build_sanitized(verbose: bool = True) -> str:
    """libbacon_gpu_san.so: api.hip's host code under ASan/UBSan, linked with
    the other translation units of the regular build (run build() first).
    Load it with BACON_GPU_LIB and the sanitizer runtime preloaded."""
    os.makedirs(OBJ, exist_ok=True)
    api_obj = os.path.join(OBJ, "api_san.o")
    cmd = [hipcc(), *HIPFLAGS, *SAN_HOST_FLAGS, "-c", os.path.join(CSRC, "api.hip"), "-o", api_obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for api.hip (sanitized):\n{r.stdout}\n{r.stderr}")
    objs = [api_obj] + [os.path.join(OBJ, os.path.splitext(s)[0] + ".o")
                        for s in SOURCES if s != "api.hip"]
    for o in objs[1:]:
        if not os.path.exists(o):
            raise RuntimeError(f"{o} missing: run build() first")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-shared-libasan",
           "-fsanitize=address,undefined", "-o", SAN_LIB, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed (sanitized):\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {SAN_LIB}")
    return SAN_LIB


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_cpp_tests(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
