"""Shared pieces of the batched-circuit tests: random batch circuits and their
CPU simulation through the oracle."""
from __future__ import annotations

import numpy as np

from qulacs_amd.batch import pauli_masks
from tests.golden.gen import haar_unitary


def random_batch_circuit(rng, n, count, max_dense=4):
    gates = []
    for _ in range(count):
        r = int(rng.integers(6))
        if r <= 1 or n < 2:
            k = 1 if n < 2 else int(rng.integers(1, min(max_dense, n) + 1))
            q = [int(x) for x in rng.choice(n, k, replace=False)]
            gates.append(("dense", q, haar_unitary(rng, k)))
        elif r == 2:
            k = int(rng.integers(1, min(4, n) + 1))
            q = [int(x) for x in rng.choice(n, k, replace=False)]
            gates.append(("diag", q, np.exp(1j * rng.uniform(0, 2 * np.pi, 1 << k))))
        elif r == 3:
            gates.append(("CZ", [int(x) for x in rng.choice(n, 2, replace=False)], None))
        elif r == 4:
            gates.append(("CNOT", [int(x) for x in rng.choice(n, 2, replace=False)], None))
        else:
            q = int(rng.integers(n))
            gates.append(("dense", [q], haar_unitary(rng, 1)))
    return gates


def random_terms(rng, n, count):
    terms = []
    for _ in range(count):
        p = rng.integers(0, 4, size=n)
        t = [q for q in range(n) if p[q]]
        bf, pf, rot = pauli_masks(t, [int(p[q]) for q in t])
        terms.append((bf, pf, rot, float(rng.standard_normal())))
    return terms


def masks_to_lists(bf, pf, n):
    t, p = [], []
    for q in range(n):
        x, z = (bf >> q) & 1, (pf >> q) & 1
        if x or z:
            t.append(q)
            p.append(2 if (x and z) else (1 if x else 3))
    return t, p


def oracle_run(oracle, n, gates, terms):
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    for kind, q, m in gates:
        if kind == "dense":
            oracle.dense(q, m, psi)
        elif kind == "diag":
            oracle.controlled_diag([], [], q, m, psi)
        else:
            oracle.two(kind, q[0], q[1], psi)
    e = 0.0
    for bf, pf, rot, c in terms:
        t, p = masks_to_lists(bf, pf, n)
        e += c * oracle.expectation(t, p, psi)
    return psi, e
