"""Builds libssw.so (the C-ABI solver library) in-tree with hipcc for gfx950.

    python build.py            # compiles csrc/*.hip and csrc/*.cpp, links libssw.so
    python build.py --checked  # the same sources with -DSSW_CHECKED: libssw_checked.so, whose
                               # kernels check every LDS address and HBM buffer index they form
                               # and whose solves fail with SSW_ERR_INTERNAL naming the first
                               # one out of bounds (a debugging build, slower than libssw.so)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libssw.so")
OBJ_CHECKED = os.path.join(HERE, "_build_checked")
LIB_CHECKED = os.path.join(HERE, "libssw_checked.so")
ARCH = os.environ.get("SSW_OFFLOAD_ARCH", "gfx950")

SOURCES = ["ssw_batch.hip", "ssw_large.hip", "ssw_probe.hip", "ssw_runtime.cpp", "ssw_large.cpp", "ssw_gen.cpp",
           "ssw_server.cpp", "ssw_server_core.cpp", "ssw_tune.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]
# packed FP32 instructions are kept out of every code object
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str, checked: bool = False) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ_CHECKED if checked else OBJ, src + ".o")
    extra = ["-DSSW_CHECKED"] if checked else []
    cmd = [hipcc(), *COMMON, *extra, *DEVICE, "-c", path, "-o", obj]
    if src.endswith(".cpp"):
        # host code: no device code object other than gfx950's (empty)
        cmd = [hipcc(), *COMMON, *extra, f"--offload-arch={ARCH}", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, checked: bool = False) -> str:
    lib = LIB_CHECKED if checked else LIB
    os.makedirs(OBJ_CHECKED if checked else OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, checked), srcs))
    tmp = lib + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", *DEVICE, "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build(verbose=True, checked="--checked" in sys.argv[1:])
    sys.exit(0)
