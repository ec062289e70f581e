"""Compile the HIP library (libqmcb_amd.so) in-tree for gfx950 with hipcc."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libqmcb_amd.so")
SOURCES = ["basis.hip", "orbital_gemm.hip", "slater.hip", "jastrow.hip", "cusp.hip", "multidet.hip",
           "walk_ae.hip", "qmcb_abi.cpp", "qmcb_stats.cpp", "qmcb_cusp.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-pass-failed"]


def _compile(src: str, obj_dir: str = OBJ, extra=()) -> str:
    obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False, lib_path: str = LIB, obj_dir: str = OBJ, extra=(),
          example: bool = True) -> str:
    os.makedirs(obj_dir, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, obj_dir, extra), SOURCES))
    tmp = lib_path + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, lib_path)
    if not example:
        return lib_path
    # the C++ drop-in example over include/qmcb_functions.hpp
    ex = os.path.join(HERE, "examples")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(os.path.dirname(HERE), "include"),
                        os.path.join(ex, "dropin_vmc.cpp"), "-L", HERE, "-lqmcb_amd",
                        "-Wl,-rpath,$ORIGIN/..", "-o", os.path.join(ex, "dropin_vmc")],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("drop-in example failed to build:\n" + r.stderr)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
