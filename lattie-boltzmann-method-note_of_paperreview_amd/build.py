"""Build the gfx950 solver library in-tree: lib/libislbm_amd.so.

    python lattie-boltzmann-method-note_of_paperreview_amd/build.py

Every HIP translation unit is compiled for gfx950 only, with packed-FP32 instruction
selection disabled, then linked into one shared library with a plain C ABI
(include/islbm.h).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libislbm_amd.so")

SOURCES = ["islbm_geometry.cpp", "islbm_kernels.hip", "islbm_tiled.hip", "islbm_api.hip", "islbm_stats.hip",
           "islbm_serve.cpp", "islbm_io.cpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"]
    cmd = [HIPCC, *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(len(srcs), 6)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_driver(verbose)
    return LIB


ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
DRIVER = os.path.join(LIB_DIR, "islbm_run")


def build_driver(verbose: bool = False) -> str:
    """The C time-loop driver (csrc/islbm_run.c): plain C, host compiler, linked against the
    library and the HIP runtime, found next to it at run time."""
    tmp = DRIVER + ".tmp"
    cmd = [os.environ.get("CC", "gcc"), "-O2", "-std=c11", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROCM, "include"), os.path.join(CSRC, "islbm_run.c"), "-o", tmp,
           "-L", LIB_DIR, "-lislbm_amd", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 or r.stderr.strip():
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, DRIVER)
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
