"""Builds libcrvpt.so (the gfx950 path tracer behind include/crvpt.h) in-tree with hipcc.

Every HIP translation unit is compiled for gfx950 only, with packed-FP32 instruction selection disabled; host-only
C++ units are compiled by hipcc as plain C++17. Objects go to cloudrendering_amd/_build/, the shared library to
cloudrendering_amd/libcrvpt.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
BUILD_DIR = os.environ.get("CRVPT_BUILD_DIR", os.path.join(PKG_DIR, "_build"))
LIB_PATH = os.environ.get("CRVPT_LIBRARY", os.path.join(PKG_DIR, "libcrvpt.so"))

HIP_SOURCES = ["render.hip", "batch.hip", "denoise.hip", "kpn.hip", "kpn_net.hip"]
CXX_SOURCES = ["api.cpp", "nanovdb_build.cpp", "kpn_model.cpp"]

ARCH = os.environ.get("CRVPT_OFFLOAD_ARCH", "gfx950")
EXTRA_HIP_FLAGS = os.environ.get("CRVPT_EXTRA_HIP_FLAGS", "").split()


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the gfx950 library cannot be built")
    return path


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    common = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-I", CSRC, "-I", os.path.join(REPO_DIR, "include")]
    if src.endswith(".hip"):
        cmd = [hipcc(), "-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", "-munsafe-fp-atomics", *EXTRA_HIP_FLAGS, *common, "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        # host-only units: plain C++ against the HIP runtime headers (no device compilation)
        rocm = os.path.dirname(os.path.dirname(os.path.realpath(hipcc())))
        cmd = [hipcc(), "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(rocm, "include"), *common, "-c",
               os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES + CXX_SOURCES))
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
    return LIB_PATH


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
