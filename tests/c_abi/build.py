"""Builds tests/c_abi/_build/step_harness: the plain-C caller of the C ABI, compiled with gcc
against include/gsh.h and linked to the in-tree gschism_amd/libgsh.so (found at run time
through a relative rpath, so the tree can move)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
OUT = HERE / "_build" / "step_harness"


def build() -> Path:
    lib = ROOT / "gschism_amd" / "libgsh.so"
    if not lib.exists():
        raise RuntimeError(f"{lib} is not built")
    OUT.parent.mkdir(exist_ok=True)
    cc = shutil.which("gcc") or "gcc"
    cmd = [cc, "-O2", "-std=c11", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}", str(HERE / "step_harness.c"),
           f"-L{lib.parent}", "-l:libgsh.so", "-Wl,-rpath,$ORIGIN/../../../gschism_amd", "-lm", "-o", str(OUT)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"harness build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return OUT


if __name__ == "__main__":
    print(build())
