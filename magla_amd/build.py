"""Build libmagla.so (host C++ + gfx950 HIP kernels) in-tree with hipcc.

    python -m magla_amd.build            # or magla_amd.build.build()

Objects go to magla_amd/_build/, the library to magla_amd/libmagla.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmagla.so")

SOURCES = ["symbolic.cpp", "csr.cpp", "capi.cpp", "factor.hip", "spmv.hip", "mpnn.hip", "bicgstab.hip"]
ARCH = os.environ.get("MAGLA_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _flags(src):
    common = ["-O3", "-fPIC", "-std=c++17", "-fwrapv", "-Wall", "-Wno-unused-function",
              "-Wno-unused-result", f"--offload-arch={ARCH}", "-I", os.path.join(HERE, "..", "include")]
    if src.endswith(".hip"):
        # packed FP32 instructions are kept out of every kernel
        common += ["-x", "hip", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                   "-munsafe-fp-atomics"]
    return common


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + _flags(src) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, err in results:
        if verbose and err.strip():
            print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
