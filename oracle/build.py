"""Builds the CPU oracle (test infrastructure: parity checker and CPU baseline).

    python oracle/build.py        ->  oracle/_build/liboracle.so

gcc, -O3, OpenMP, no floating-point contraction and no fast-math, so the oracle evaluates
every expression in the written order.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build"
LIB = OUT / "liboracle.so"
SRC = HERE / "gsh_oracle.c"
HDR = HERE / "gsh_oracle.h"
CC = shutil.which("gcc") or "gcc"
FLAGS = ["-O3", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
         "-march=x86-64-v2", "-Wall", "-Wno-unused-function"]


def build(force: bool = False) -> Path:
    OUT.mkdir(exist_ok=True)
    if force or not LIB.exists() or max(SRC.stat().st_mtime, HDR.stat().st_mtime) > LIB.stat().st_mtime:
        tmp = LIB.with_suffix(".so.tmp")
        r = subprocess.run([CC, *FLAGS, str(SRC), "-o", str(tmp), "-lm"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, LIB)
    return LIB


def build_planted(k: int) -> Path:
    """The oracle with deliberate mistake k (OR_PLANT in gsh_oracle.c) built into its own library,
    for the tests that show a pin fails on it."""
    OUT.mkdir(exist_ok=True)
    lib = OUT / f"liboracle_plant{k}.so"
    if not lib.exists() or max(SRC.stat().st_mtime, HDR.stat().st_mtime) > lib.stat().st_mtime:
        tmp = lib.with_suffix(".so.tmp")
        r = subprocess.run([CC, *FLAGS, f"-DOR_PLANT={int(k)}", str(SRC), "-o", str(tmp), "-lm"], capture_output=True,
                           text=True)
        if r.returncode != 0:
            raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
