"""Build recipe for libgds2para.so (gfx950) and the CPU oracle.

Every HIP translation unit is compiled by hipcc for gfx950 only, with packed
FP32 instructions disabled; the exact-arithmetic kernels (stiffness assembly,
V0 basis) are additionally compiled with FP contraction off so their products
and sums round as written.  Objects go to gds2para_amd/_build/, the shared
library to gds2para_amd/lib/ (in-tree, so it travels with the repository
snapshot).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-std=c++17", "-O3", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-result"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]
EXACT = {"assemble.hip", "v0.hip", "host_mesh.cpp"}


def _sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".hip", ".cpp")):
            out.append(f)
    return out


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src + ".o")
    path = os.path.join(CSRC, src)
    flags = list(COMMON)
    if src.endswith(".hip"):
        flags += ["-x", "hip"] + HIP_FLAGS
    if src in EXACT:
        flags += ["-ffp-contract=off"]
    compiler = HIPCC if src.endswith(".hip") else os.environ.get("CXX", "g++")
    cmd = [compiler] + flags + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    lib = os.path.join(LIBDIR, "libgds2para.so")
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


def build_oracle(verbose: bool = False) -> str:
    d = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-C", d], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(r.stdout.strip())
    return os.path.join(d, "_build", "libgp_oracle.so")


def build(verbose: bool = False):
    build_library(verbose)


if __name__ == "__main__":
    build_library(verbose=True)
    if "--oracle" in sys.argv:
        build_oracle(verbose=True)
