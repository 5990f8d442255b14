"""Build libtofe.so (HIP kernels + C ABI) in-tree for gfx950.

Every translation unit under csrc/ is compiled with hipcc in parallel into
topopt3d_amd/_build/, then linked into topopt3d_amd/libtofe.so against the HIP runtime only.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtofe.so")
ARCH = os.environ.get("TOFE_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I" + CSRC, "-I" + os.path.join(PKG, "..", "include")]
# Device code: gfx950 only; packed FP32 instructions are kept out of the generated code.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += DEVICE
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=1200)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + [
        "-pthread", "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_examples()
    if verbose:
        print("built", LIB)
    return LIB


EXAMPLES = os.path.join(PKG, "examples")


def build_examples() -> None:
    """C++ host programs over the C ABI (include/tofe_topopt.hpp), linked to libtofe.so."""
    cxx = shutil.which("g++") or "g++"
    for src in sorted(os.listdir(EXAMPLES)):
        if not src.endswith(".cpp"):
            continue
        out = os.path.join(EXAMPLES, src[:-4])
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(PKG, "..", "include"),
               os.path.join(EXAMPLES, src), "-o", out, "-L" + PKG, "-ltofe",
               "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"example build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
