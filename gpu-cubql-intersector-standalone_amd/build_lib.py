"""Build recipe for libgi.so (HIP kernels for gfx950 + the C ABI).

Compiles every source under csrc/ with hipcc and links them in-tree into
libgi.so next to this file. Flags:
  --offload-arch=gfx950      MI355X only
  -ffp-contract=off          the FP64 filter's error bound and the CGAL-order
                             transform need separately rounded operations
  -target-feature -packed-fp32-ops
                             keep the compiler from emitting packed FP32 ops
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgi.so")
ROOT = os.path.dirname(HERE)

SOURCES = ["gi_device.hip", "gi_api.hip", "gi_noalloc.hip", "gi_serve.cpp", "gi_meshio.cpp", "host_exact.cpp", "../host/KernelBVHController.cpp"]
FACADE_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "facade_test.cpp")
FACADE_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "facade_test")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(os.path.basename(src))[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build() -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libgi.so failed")
    os.replace(tmp, LIB)
    return LIB


def build_facade_test() -> str:
    """The C++ driver of the KernelBVHController facade used by the GPU tests
    (links libgi.so from this directory)."""
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(HERE, "host"), FACADE_TEST_SRC, "-o", FACADE_TEST_BIN,
           "-L" + HERE, "-lgi", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../../gpu-cubql-intersector-standalone_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("facade test driver failed to build")
    return FACADE_TEST_BIN


if __name__ == "__main__":
    print(build())
