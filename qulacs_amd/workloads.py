"""Seeded synthetic workloads for the gate-sweep benchmark and its tests.

The single-large-state sweep: an n-qubit complex128 state and `count` gates
drawn uniformly from five families --
  (a) a 1-qubit Haar-random dense unitary on a uniform target,
  (b) a 2-qubit Haar-random 4x4 dense unitary on a random pair,
  (c) RX / RY / RZ with angle U[0, 2 pi),
  (d) CZ / CNOT on a random pair,
  (e) a 5-qubit dense 32x32 unitary formed by fusing 1- and 2-qubit Haar gates
      on random targets --
plus the density-matrix variant: the same circuit on a smaller density matrix
with single-qubit depolarizing noise on every touched qubit after each gate.

Haar unitaries are the Q of the QR decomposition of a complex Gaussian matrix
with the phases of R's diagonal divided out (as gate::RandomUnitary,
src/cppsim/gate_factory.cpp:341-349 in the reference).

Algorithmic bytes per gate are 2 x 16 B x (amplitudes the gate can change):
dense, RX, RY, RZ: all 2^n; CNOT: 2^(n-1); CZ: 2^(n-2).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Gate:
    kind: str                      # dense1 dense2 dense5 RX RY RZ CNOT CZ
    targets: list                  # for CNOT/CZ: [control, target]
    matrix: np.ndarray | None = None
    angle: float = 0.0
    family: str = ""               # a..e
    meta: dict = field(default_factory=dict)

    @property
    def position(self) -> str:
        """'low' if every acted-on qubit is below 6 (inside one 64-amplitude run),
        'high' if all are >= 6, else 'mixed'."""
        lo = [t < 6 for t in self.targets]
        return "low" if all(lo) else ("high" if not any(lo) else "mixed")


def haar_unitary(rng: np.random.Generator, k: int) -> np.ndarray:
    d = 1 << k
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return np.ascontiguousarray(q * ph[None, :])


def _embed(k: int, local_targets, u: np.ndarray) -> np.ndarray:
    """2^k x 2^k operator of `u` acting on local qubits `local_targets` (bit i <-> local_targets[i])."""
    d = 1 << k
    full = np.zeros((d, d), dtype=np.complex128)
    tmask = 0
    for t in local_targets:
        tmask |= 1 << t
    for col in range(d):
        x = 0
        for i, t in enumerate(local_targets):
            x |= ((col >> t) & 1) << i
        base = col & ~tmask
        for y in range(u.shape[0]):
            row = base
            for i, t in enumerate(local_targets):
                row |= ((y >> i) & 1) << t
            full[row, col] += u[y, x]
    return full


def fused_unitary(rng: np.random.Generator, k: int = 5, parts: int = 8) -> np.ndarray:
    """Product of `parts` random 1- and 2-qubit Haar gates on k local qubits."""
    m = np.eye(1 << k, dtype=np.complex128)
    for _ in range(parts):
        if rng.random() < 0.5:
            t = [int(rng.integers(k))]
        else:
            t = rng.choice(k, 2, replace=False).tolist()
        m = _embed(k, t, haar_unitary(rng, len(t))) @ m
    return m


def gate_sweep(n: int = 30, count: int = 200, seed: int = 2002) -> list[Gate]:
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(count):
        fam = "abcde"[int(rng.integers(5))]
        if fam == "a":
            gates.append(Gate("dense1", [int(rng.integers(n))], haar_unitary(rng, 1), family=fam))
        elif fam == "b":
            gates.append(Gate("dense2", rng.choice(n, 2, replace=False).tolist(), haar_unitary(rng, 2), family=fam))
        elif fam == "c":
            kind = ("RX", "RY", "RZ")[int(rng.integers(3))]
            gates.append(Gate(kind, [int(rng.integers(n))], angle=float(rng.uniform(0, 2 * np.pi)), family=fam))
        elif fam == "d":
            kind = ("CZ", "CNOT")[int(rng.integers(2))]
            gates.append(Gate(kind, rng.choice(n, 2, replace=False).tolist(), family=fam))
        else:
            gates.append(Gate("dense5", rng.choice(n, 5, replace=False).tolist(), fused_unitary(rng), family=fam))
    return gates


def alg_bytes(g: Gate, n: int) -> int:
    """Algorithmic HBM bytes of one gate on an n-qubit state (2 x 16 B x changed amplitudes)."""
    if g.kind == "CNOT":
        return 32 << (n - 1)
    if g.kind == "CZ":
        return 32 << (n - 2)
    return 32 << n


def rotation_matrix(kind: str, angle: float) -> np.ndarray:
    """qulacs convention exp(+i angle P / 2) (src/csim/update_ops_pauli_single.cpp:54-116)."""
    c, s = np.cos(angle / 2), np.sin(angle / 2)
    if kind == "RX":
        return np.array([[c, 1j * s], [1j * s, c]])
    if kind == "RY":
        return np.array([[c, s], [-s, c]])
    return np.diag([c + 1j * s, c - 1j * s])


# ---------------------------------------------------------------- driving the C ABI
def _cmat(lib_mod, m: np.ndarray):
    a = np.ascontiguousarray(m, dtype=np.complex128).ravel()
    arr = (lib_mod.QmComplex * a.size)()
    C.memmove(arr, a.ctypes.data, a.nbytes)
    return arr


class GateLauncher:
    """Pre-marshals a gate list for the C ABI so the timed loop only makes calls."""

    def __init__(self, lib_mod, gates: list[Gate]):
        self.m = lib_mod
        self.calls = []
        L = lib_mod.lib()
        for g in gates:
            if g.kind == "dense1":
                self.calls.append((L.single_qubit_dense_matrix_gate_host, (g.targets[0], _cmat(lib_mod, g.matrix))))
            elif g.kind == "dense2":
                self.calls.append((L.double_qubit_dense_matrix_gate_host,
                                   (g.targets[0], g.targets[1], _cmat(lib_mod, g.matrix))))
            elif g.kind == "dense5":
                tg = (C.c_uint * 5)(*g.targets)
                self.calls.append((L.penta_qubit_dense_matrix_gate_host, (tg, _cmat(lib_mod, g.matrix))))
            elif g.kind in ("RX", "RY", "RZ"):
                self.calls.append((getattr(L, f"{g.kind}_gate_host"), (g.targets[0], g.angle)))
            elif g.kind in ("CNOT", "CZ"):
                self.calls.append((getattr(L, f"{g.kind}_gate_host"), (g.targets[0], g.targets[1])))
            else:
                raise ValueError(g.kind)

    def launch(self, i: int, state_ptr, dim: int, stream, device: int = 0) -> None:
        fn, args = self.calls[i]
        fn(*args, state_ptr, dim, stream, device)


# ---------------------------------------------------------------- many small circuits
@dataclass
class SmallCircuit:
    """One circuit of the many-small-circuits workload: gates in batch form
    (kind, qubits, matrix) and the observable as (bit_flip, phase_flip, rot90, coef)."""

    n: int
    layers: int
    gates: list
    terms: list
    field: float


def tfim_terms(n: int, h: float) -> list:
    """H = sum_{i<n-1} Z_i Z_{i+1} + h sum_i X_i in mask form."""
    terms = [(0, (1 << i) | (1 << (i + 1)), 0, 1.0) for i in range(n - 1)]
    terms += [(1 << i, 0, 0, h) for i in range(n)]
    return terms


def many_small_circuits(count: int = 4096, seed: int = 3003, n_range=(8, 16), layer_range=(4, 12)) -> list:
    """count circuits with n ~ U{n_range}, L ~ U{layer_range} layers; a layer is
    RY(theta) then RZ(phi) on every qubit (angles U[0, 2 pi)), then CZ(i, i+1)
    for i = 0..n-2; observable: transverse-field Ising H with h ~ U[0.5, 1.5]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        L = int(rng.integers(layer_range[0], layer_range[1] + 1))
        gates = []
        for _ in range(L):
            th = rng.uniform(0, 2 * np.pi, n)
            ph = rng.uniform(0, 2 * np.pi, n)
            for q in range(n):
                gates.append(("dense", [q], rotation_matrix("RY", th[q])))
                gates.append(("dense", [q], rotation_matrix("RZ", ph[q])))
            for q in range(n - 1):
                gates.append(("CZ", [q, q + 1], None))
        h = float(rng.uniform(0.5, 1.5))
        out.append(SmallCircuit(n, L, gates, tfim_terms(n, h), h))
    return out


def small_circuit_alg_flops(c: SmallCircuit) -> float:
    """FP64 flops of the fused form: one 2x2 complex matrix per qubit per layer
    (8 real multiply-adds = 16 flops per amplitude)."""
    return 16.0 * c.layers * c.n * (1 << c.n)


# ---------------------------------------------------------------- 20-qubit brickwork
def brickwork_circuit(n: int = 20, layers: int = 10, terms: int = 16, samples: int = 1024, seed: int = 1001):
    """n-qubit |0> circuit of `layers` brickwork layers: a Haar 2x2 on every qubit,
    then CNOT(i, i+1) for even i (even layers) or odd i (odd layers); observable of
    `terms` Pauli strings (uniform I/X/Y/Z per qubit, coef ~ N(0, 1)) and
    `samples` uniforms in [0, 1).  Gates are (kind, qubits, matrix) with kinds
    "dense" and "CNOT" (control first); terms are (bit_flip, phase_flip, rot90, coef)."""
    rng = np.random.default_rng(seed)
    gates = []
    for layer in range(layers):
        for q in range(n):
            gates.append(("dense", [q], haar_unitary(rng, 1)))
        for q in range(layer % 2, n - 1, 2):
            gates.append(("CNOT", [q, q + 1], None))
    obs = []
    for _ in range(terms):
        p = rng.integers(0, 4, size=n)
        bf = pf = rot = 0
        for q in range(n):
            if p[q] in (1, 2):
                bf |= 1 << q
            if p[q] in (2, 3):
                pf |= 1 << q
            rot += int(p[q] == 2)
        obs.append((bf, pf, rot, float(rng.standard_normal())))
    return gates, obs, rng.random(samples)


# ---------------------------------------------------------------- serving stream
@dataclass
class ServingRequest:
    arrival: float
    n: int
    depth: int
    gates: list
    terms: list
    uniforms: np.ndarray


def serving_requests(count: int = 64, rate: float = 50.0, seed: int = 5005, n_range=(10, 28),
                     depth_range=(20, 200), terms: int = 8, samples: int = 1000) -> list:
    """Poisson arrivals at `rate` per second; each request is a brickwork circuit
    (a Haar 2x2 on every qubit, then CNOT(i, i+1) on even/odd i by layer) with
    n = round(exp(U[ln lo, ln hi])) qubits and depth ~ U{depth_range}, a
    `terms`-term random Pauli observable (coef ~ N(0, 1)) and `samples` uniforms."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += float(rng.exponential(1.0 / rate))
        n = int(round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        gates = []
        for layer in range(depth):
            for q in range(n):
                gates.append(("dense", [q], haar_unitary(rng, 1)))
            for q in range(layer % 2, n - 1, 2):
                gates.append(("CNOT", [q, q + 1], None))
        obs = []
        for _ in range(terms):
            p = rng.integers(0, 4, size=n)
            bf = pf = rot = 0
            for q in range(n):
                if p[q] in (1, 2):
                    bf |= 1 << q
                if p[q] in (2, 3):
                    pf |= 1 << q
                rot += int(p[q] == 2)
            obs.append((bf, pf, rot, float(rng.standard_normal())))
        out.append(ServingRequest(t, n, depth, gates, obs, rng.random(samples)))
    return out
