"""Builds libqulacs_mi355.so (HIP kernels + C ABI) for gfx950 and the Python
module qulacs_amd._core over it, in-tree.

Every translation unit under qulacs_amd/csrc is compiled with hipcc (``.hip``)
or g++ (``.cpp``) into qulacs_amd/_build/, then linked into
qulacs_amd/libqulacs_mi355.so.  The host-side qulacs API (csrc/host: C++ classes
over the C ABI plus their pybind11 bindings) is linked into
qulacs_amd/_core<ext-suffix>.so, which finds the library next to itself through
an $ORIGIN run path.  The build always recompiles; it never falls back to a
prebuilt or CPU library.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libqulacs_mi355.so"
HOST = CSRC / "host"
CORE = PKG / ("_core" + sysconfig.get_config_var("EXT_SUFFIX"))
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
# Per-file device flags.  qm_batch.hip and qm_adjoint.hip dispatch register-array
# updates through a switch on a runtime bit position; LLVM's common-code sinking
# would merge the cases into one access with a computed index and move the array
# to private (scratch) memory, so it is switched off for those files.
EXTRA_HIP_FLAGS = {"qm_batch.hip": ["-mllvm", "-simplifycfg-sink-common=false"],
                   "qm_adjoint.hip": ["-mllvm", "-simplifycfg-sink-common=false"]}
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


def _pybind_includes() -> list[str]:
    import pybind11

    return [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every source, link the shared library and the Python module.
    Returns the library's path."""
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
            cmds.append([HIPCC, *HIP_FLAGS, *EXTRA_HIP_FLAGS.get(src.name, []), "-c", str(src), "-o", str(obj)])
        else:
            cmds.append([CXX, *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
    host_objs = []
    for src in sorted(HOST.glob("*.cpp")):
        obj = BUILD / ("host_" + src.name + ".o")
        host_objs.append(obj)
        cmds.append([CXX, *CXX_FLAGS, *_pybind_includes(), "-c", str(src), "-o", str(obj)])
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)], verbose)
    os.replace(tmp, LIB)
    tmp = CORE.with_suffix(".tmp")
    _run([CXX, "-shared", "-fPIC", "-o", str(tmp), *map(str, host_objs), f"-L{PKG}", "-lqulacs_mi355",
          "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined", f"-L{sysconfig.get_config_var('LIBDIR')}",
          f"-lpython{sysconfig.get_config_var('LDVERSION')}"], verbose)
    os.replace(tmp, CORE)
    return LIB


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
