"""The drop-in boundary exercised as cppsim exercises it: a C++ program written
against the reference's seam declarations (restated in tests/seam/seam_decls.hpp,
C++ linkage, so it references the mangled names) is compiled here with g++,
linked against libqulacs_mi355.so, and run; its final state and scalars equal
the oracle's replay of the same gate list."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.golden.gen import haar_unitary, random_state  # noqa: F401

ROOT = Path(__file__).resolve().parents[1]
SEAM = ROOT / "tests" / "seam"
LIB = ROOT / "qulacs_amd" / "libqulacs_mi355.so"
SYMS = ROOT / "tests" / "golden" / "reference_seam_symbols.txt"


def _build(tmp_path: Path) -> Path:
    exe = tmp_path / "seam_driver"
    subprocess.run(["g++", "-std=c++14", "-O1", f"-I{SEAM}", str(SEAM / "seam_driver.cpp"), str(LIB),
                    f"-Wl,-rpath,{LIB.parent}", "-o", str(exe)], check=True)
    return exe


def test_seam_declarations_mangle_like_the_reference(tmp_path):
    """Every function the driver calls is one of the reference seam's mangled
    names (so a program built on the reference headers binds the same symbols)."""
    obj = tmp_path / "d.o"
    subprocess.run(["g++", "-std=c++14", "-c", f"-I{SEAM}", str(SEAM / "seam_driver.cpp"), "-o", str(obj)],
                   check=True)
    nm = subprocess.run(["nm", "-u", str(obj)], check=True, capture_output=True, text=True).stdout
    used = {l.split()[-1] for l in nm.splitlines() if "_host" in l or "get_num_device" in l}
    ref = {l.strip() for l in SYMS.read_text().splitlines() if l.strip()}
    assert len(used) >= 25 and used <= ref, sorted(used - ref)


def _fmt(m) -> str:
    a = np.asarray(m, dtype=np.complex128).ravel()
    return " ".join(f"{z.real:.17g} {z.imag:.17g}" for z in a)


@pytest.mark.gpu
def test_mangled_seam_program_matches_oracle(oracle, tmp_path):
    exe = _build(tmp_path)
    rng = np.random.default_rng(91)
    n = 12
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    lines = [f"n {n}"]
    named = {"h": np.array([[1, 1], [1, -1]]) / np.sqrt(2), "s": np.diag([1, 1j]),
             "t": np.diag([1, np.exp(1j * np.pi / 4)]), "sqrtx": np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]) / 2}
    scalars = []
    for i in range(60):
        k = i % 10
        if k == 0:
            t = int(rng.integers(n)); u = haar_unitary(rng, 1)
            lines.append(f"d1 {t} {_fmt(u)}"); oracle.dense1(t, u, psi)
        elif k == 1:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False)); u = haar_unitary(rng, 2)
            lines.append(f"d2 {a} {b} {_fmt(u)}"); oracle.dense2(a, b, u, psi)
        elif k == 2:
            tg = [int(x) for x in rng.choice(n, 3, replace=False)]; u = haar_unitary(rng, 3)
            lines.append(f"dk 3 {' '.join(map(str, tg))} {_fmt(u)}"); oracle.dense(tg, u, psi)
        elif k == 3:
            c, t = (int(x) for x in rng.choice(n, 2, replace=False)); u = haar_unitary(rng, 1); v = int(rng.integers(2))
            lines.append(f"cd {c} {v} {t} {_fmt(u)}"); oracle.controlled_dense([c], [v], [t], u, psi)
        elif k == 4:
            t = int(rng.integers(n)); d = np.exp(1j * rng.uniform(0, 6, 2))
            lines.append(f"diag1 {t} {_fmt(d)}"); oracle.diag1(t, d, psi)
        elif k == 5:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False)); name = ("cnot", "cz", "swap")[i % 3]
            lines.append(f"{name} {a} {b}"); oracle.two({"cnot": "CNOT", "cz": "CZ", "swap": "SWAP"}[name], a, b, psi)
        elif k == 6:
            t = int(rng.integers(n)); ang = float(rng.uniform(0, 6.3)); name = ("rx", "ry", "rz")[i % 3]
            # RX/RY/RZ(a) = exp(+i a P / 2) (src/csim/update_ops_pauli_single.cpp:54-116)
            lines.append(f"{name} {t} {ang!r}"); oracle.pauli_rotation([t], [{"rx": 1, "ry": 2, "rz": 3}[name]], ang, psi)
        elif k == 7:
            t = int(rng.integers(n)); name = ("h", "s", "t", "sqrtx")[i % 4]
            lines.append(f"{name} {t}"); oracle.dense1(t, named[name], psi)
        elif k == 8:
            tg = [int(x) for x in rng.choice(n, 3, replace=False)]; pl = [int(x) for x in rng.integers(1, 4, 3)]
            ang = float(rng.uniform(0, 6.3))
            lines.append(f"prot 3 {' '.join(map(str, tg))} {' '.join(map(str, pl))} {ang!r}")
            oracle.pauli_rotation(tg, pl, ang, psi)
        else:
            tg = [int(x) for x in rng.choice(n, 4, replace=False)]; pl = [int(x) for x in rng.integers(1, 4, 4)]
            lines.append(f"exp 4 {' '.join(map(str, tg))} {' '.join(map(str, pl))}")
            scalars.append(("exp", oracle.expectation(tg, pl, psi)))
    lines += ["norm", "m0 3", "inner"]
    scalars += [("norm", oracle.norm_squared(psi)), ("m0", oracle.M0(3, psi)), ("inner", oracle.norm_squared(psi))]
    ops = tmp_path / "ops.txt"
    ops.write_text("\n".join(lines) + "\n")
    out = tmp_path / "state.bin"
    print("running the seam driver", flush=True)
    r = subprocess.run([str(exe), str(ops), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, dtype=np.complex128)
    assert np.linalg.norm(got - psi) / np.linalg.norm(psi) < 1e-10
    rows = [l.split() for l in r.stdout.splitlines() if l.strip()]
    assert [x[0] for x in rows] == [s[0] for s in scalars]
    for row, (_, want) in zip(rows, scalars):
        assert abs(float(row[1]) - want) < 1e-10, (row, want)
