"""Builds libqulacs_mi355.so (HIP kernels + C ABI) for gfx950, in-tree.

Every translation unit under qulacs_amd/csrc is compiled with hipcc (``.hip``)
or g++ (``.cpp``) into qulacs_amd/_build/, then linked into
qulacs_amd/libqulacs_mi355.so, which the package loads with ctypes.  The build
always recompiles; it never falls back to a prebuilt or CPU library.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libqulacs_mi355.so"
ROOT = PKG.parent

ARCH = os.environ.get("QM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep hipcc from forming packed FP32 instructions (device pass only; the
    # host pass warns that it does not know the feature)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unknown-warning-option",
    f"-I{ROOT / 'include'}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-Wall", f"-I{ROOT / 'include'}"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    if verbose and out.strip():
        print(out, flush=True)


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every source and link the shared library.  Returns its path."""
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    if not srcs:
        raise RuntimeError(f"no sources under {CSRC}")
    objs = []
    cmds = []
    for src in srcs:
        obj = BUILD / (src.name + ".o")
        objs.append(obj)
        if src.suffix == ".hip":
            cmds.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        else:
            cmds.append([CXX, *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
