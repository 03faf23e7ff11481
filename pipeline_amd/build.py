"""Builds pipeline_amd/_lib/libocpipe.so from pipeline_amd/csrc with hipcc
for gfx950 (one object per stage, compiled in parallel, then linked)."""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
BUILD_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libocpipe.so")

SOURCES = ["k_util.hip", "k_init.hip", "k_morton.hip", "k_onesweep.hip", "k_unique.hip",
           "k_radix_tree.hip", "k_edge_scan.hip", "k_octree.hip", "k_octree_stage.hip",
           "k_learned.hip", "k_msd.hip", "k_format.hip", "capi.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("OCP_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # Products and sums must round separately to match the CPU path bit for
    # bit (Morton quantisation and corner decode).
    "-ffp-contract=off",
    # Keep hipcc from emitting packed FP32 arithmetic.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_c_demo()
    build_fixture_demo()
    if verbose:
        print(f"built {LIB}")
    return LIB


ROOT = os.path.dirname(HERE)
C_DEMO_SRC = os.path.join(ROOT, "tests", "c", "abi_demo.c")
C_DEMO = os.path.join(OUT_DIR, "abi_demo")


def build_c_demo(out: str = C_DEMO) -> str:
    """Compiles the plain-C caller of include/octree_pipeline.h (C99, no C++)
    and links it against the library: the header is valid C and every entry
    point it uses links."""
    cc = os.environ.get("CC", shutil.which("gcc") or "cc")
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    cmd = [cc, "-std=c99", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           C_DEMO_SRC, "-o", out, "-L", OUT_DIR, "-locpipe", "-L", rocm_lib,
           f"-Wl,-rpath,{OUT_DIR}", f"-Wl,-rpath,{rocm_lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C demo build failed:\n{r.stdout}\n{r.stderr}")
    return out


FIXTURE_DEMO_SRC = os.path.join(ROOT, "tests", "c", "fixture_demo.cpp")
FIXTURE_DEMO = os.path.join(OUT_DIR, "fixture_demo")


def build_fixture_demo(out: str = FIXTURE_DEMO) -> str:
    """Compiles the C++ caller of include/octree_fixture.hpp with g++ (host
    code only) and links it against the library."""
    cxx = os.environ.get("CXX", shutil.which("g++") or "c++")
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I",
           os.path.join(ROOT, "include"), FIXTURE_DEMO_SRC, "-o", out, "-L", OUT_DIR,
           "-locpipe", "-L", rocm_lib, f"-Wl,-rpath,{OUT_DIR}", f"-Wl,-rpath,{rocm_lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"fixture demo build failed:\n{r.stdout}\n{r.stderr}")
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
