"""Ahead-of-time build of libgsh.so (HIP for gfx950 + host C++), in-tree.

Every source is compiled without floating-point contraction and without fast-math, so
device results follow the written operation order exactly.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libgsh.so"
INCLUDE = PKG.parent / "include"

ARCH = os.environ.get("GSH_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}"]
HIP_FLAGS = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-fno-gpu-rdc"]
HOST_SOURCES = ["gen.cpp", "topology.cpp", "ssor_layout.cpp", "grid_io.cpp"]
HIP_SOURCES = ["mesh_kernels.hip", "pcg_kernels.hip", "ssor_kernels.hip", "block_kernels.hip", "api.hip"]
HEADERS = ["host_mesh.h", "abi_util.h", "context.h", "device.h", "reduce.h", "ssor_layout.h"]


def _stale(obj: Path, src: Path) -> bool:
    if not obj.exists():
        return True
    deps = [src, INCLUDE / "gsh.h"] + [CSRC / h for h in HEADERS]
    t = obj.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build(force: bool = False, jobs: int = 4) -> Path:
    """Compiles every source (always when force=True) and links libgsh.so."""
    BUILD.mkdir(exist_ok=True)
    work = []
    for name in HOST_SOURCES:
        src, obj = CSRC / name, BUILD / (name + ".o")
        if force or _stale(obj, src):
            work.append([CXX, *COMMON, "-c", str(src), "-o", str(obj)])
    for name in HIP_SOURCES:
        src, obj = CSRC / name, BUILD / (name + ".o")
        if force or _stale(obj, src):
            work.append([HIPCC, *COMMON, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, work))
    objs = [str(BUILD / (n + ".o")) for n in HOST_SOURCES + HIP_SOURCES]
    if force or work or not LIB.exists():
        tmp = LIB.with_suffix(".so.tmp")
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp)])
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv)
    print(LIB)
