"""Writes tests/golden/reference_seam_symbols.txt: the mangled C++ names of the
reference's GPU seam as cppsim would bind them.

A tiny translation unit that includes the reference headers
src/gpusim/{update_ops_cuda,stat_ops,memory_ops,util_func}.h and takes the
address of every declared `*_host` / device function is compiled with g++; the
undefined symbols of that object are exactly the names a cppsim built with its
GPU option needs from the library.  The list is data; tests check that
libqulacs_mi355.so defines every one of them.

Run: python tests/golden/make_seam_symbols.py
"""
from __future__ import annotations

import re
import subprocess
import tempfile
from pathlib import Path

REF = Path("/root/reference/src/gpusim")
OUT = Path(__file__).resolve().parent / "reference_seam_symbols.txt"
HEADERS = ["update_ops_cuda.h", "stat_ops.h", "memory_ops.h", "util_func.h"]


def declared_names() -> list[str]:
    names = []
    for h in HEADERS:
        text = (REF / h).read_text()
        for m in re.finditer(r"DllExport\s+[\w:<>\s\*]+?\b(\w+)\s*\(", text):
            names.append(m.group(1))
    return sorted(set(names))


def main() -> None:
    names = declared_names()
    # Functions the reference declares but cppsim never calls and the GPU seam
    # leaves unimplemented in this library.
    skip = {"multi_Z_gate_host", "multipauli_get_expectation_value_host"}
    names = [n for n in names if n not in skip]
    src = "".join(f'#include "{h}"\n' for h in HEADERS)
    src += "void* qm_seam_refs[] = {\n" + "".join(f"    (void*)&{n},\n" for n in names) + "};\n"
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "seam.cpp"
        o = Path(d) / "seam.o"
        c.write_text(src)
        subprocess.run(["g++", "-std=c++14", "-c", f"-I{REF}", str(c), "-o", str(o)], check=True)
        nm = subprocess.run(["nm", "-u", str(o)], check=True, capture_output=True, text=True).stdout
    syms = sorted({line.split()[-1] for line in nm.splitlines() if line.strip()})
    syms = [s for s in syms if s.startswith("_Z")]
    OUT.write_text("\n".join(syms) + "\n")
    print(f"{len(syms)} symbols -> {OUT}")


if __name__ == "__main__":
    main()
