"""Builds libmatqr.so (the HIP/gfx950 library behind include/mat_qr.h) in-tree.

    python -m mat_amd.build      # or mat_amd.build.build()

One hipcc invocation per translation unit (object files under
mat_amd/_build/), then one shared-library link into mat_amd/lib/.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libmatqr.so")
# the C++ drop-in test of include/mat/Mat.h, linked against LIB
CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_mat_api.cpp")
CPP_TEST_BIN = os.path.join(HERE, "lib", "test_mat_api")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM_LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))), "lib")
ARCH = os.environ.get("MQR_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-munsafe-fp-atomics",
    # keep the compiler from forming packed-FP32 VALU ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    # roctx (phase ranges for rocprofv3 --marker-trace) from the ROCm install
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           f"-L{ROCM_LIB}", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM_LIB}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_cpp_test(verbose)
    return LIB


def build_cpp_test(verbose: bool = False) -> str:
    """Host-only C++ program (g++, no HIP headers) using include/mat/Mat.h."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
           CPP_TEST_SRC, "-o", CPP_TEST_BIN, "-L", os.path.dirname(LIB), "-lmatqr",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {CPP_TEST_SRC}:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {CPP_TEST_BIN}")
    return CPP_TEST_BIN


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
