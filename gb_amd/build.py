"""Build libgbgpu.so (HIP for gfx950) in place: python -m gb_amd.build"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgbgpu.so")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions are kept out of every kernel (hipcc's host pass
# warns that it does not know the feature; that warning is harmless)
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function"]

SOURCES = ["api.hip", "sparse.hip", "dense.hip", "rows.hip", "batch.hip", "dsparse.hip",
           "noalloc.hip"]
C_SOURCES = ["gb_adapter.c", "f4.c", "f4_multi.c"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".c"):
        cmd = ["gcc", "-O2", "-fPIC", "-std=gnu11", "-Wall", "-fopenmp", "-I",
               os.path.join(ROOT, "include"), "-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + FLAGS + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"compiling {src} failed")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = SOURCES + [c for c in C_SOURCES if os.path.exists(os.path.join(CSRC, c))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    # the C host code uses OpenMP through gcc's runtime
    gomp = subprocess.run(["gcc", "-print-file-name=libgomp.so"], capture_output=True,
                          text=True).stdout.strip()
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [gomp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libgbgpu.so failed")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_adapter_driver()
    return LIB


def build_adapter_driver() -> str:
    """tests/_bin/adapter_driver: calls the gb adapters like gb's F4 loop."""
    out_dir = os.path.join(ROOT, "tests", "_bin")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "adapter_driver")
    cmd = ["gcc", "-O2", "-std=gnu11", "-Wall", os.path.join(ROOT, "tests", "adapter_driver.c"),
           "-o", exe, "-L", HERE, "-lgbgpu", "-Wl,-rpath,$ORIGIN/../../gb_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building the adapter driver failed")
    return exe


if __name__ == "__main__":
    build(verbose=True)
