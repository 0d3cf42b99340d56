"""Seeded input generators shared by the fixture writer, the tests and the bench.

Conventions follow the reference: qubit q is bit q of the index; a dense
k-qubit matrix is row-major with bit i of the row/column index belonging to
targets[i]; Haar-random unitaries are the Q of a QR decomposition of a complex
Gaussian matrix with the phases of R's diagonal divided out (as
gate::RandomUnitary does, src/cppsim/gate_factory.cpp:341-349).
"""
from __future__ import annotations

import numpy as np


def haar_unitary(rng: np.random.Generator, k: int) -> np.ndarray:
    d = 1 << k
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return np.ascontiguousarray(q * ph[None, :])


def random_state(rng: np.random.Generator, n: int) -> np.ndarray:
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return np.ascontiguousarray(v / np.linalg.norm(v))


def random_pauli_terms(rng: np.random.Generator, n: int, count: int):
    """count terms, each with a uniform I/X/Y/Z on every qubit and coef ~ N(0,1).

    Returns a list of (targets, paulis, coef); identity factors are dropped
    from the lists (an all-identity term has empty lists)."""
    out = []
    for _ in range(count):
        p = rng.integers(0, 4, size=n)
        t = [q for q in range(n) if p[q] != 0]
        out.append((t, [int(p[q]) for q in t], float(rng.standard_normal())))
    return out


def pauli_masks(targets, paulis):
    """(bit_flip_mask, phase_flip_mask, Y count, pivot) as src/csim/utility.cpp:12-43."""
    bf = pf = 0
    rot = 0
    pivot = 0
    for t, p in zip(targets, paulis):
        if p == 1:
            bf ^= 1 << t
            pivot = t
        elif p == 2:
            bf ^= 1 << t
            pf ^= 1 << t
            rot += 1
            pivot = t
        elif p == 3:
            pf ^= 1 << t
    return bf, pf, rot, pivot


# ---------------------------------------------------------------- dense closed forms
I2 = np.eye(2, dtype=np.complex128)
PAULI = [I2, np.array([[0, 1], [1, 0]], dtype=np.complex128),
         np.array([[0, -1j], [1j, 0]], dtype=np.complex128), np.array([[1, 0], [0, -1]], dtype=np.complex128)]


def expand(n: int, targets, matrix) -> np.ndarray:
    """Full 2^n x 2^n operator of a k-qubit matrix on `targets` (bit i <-> targets[i]).

    Built column by column from the definition, independent of any kernel."""
    dim = 1 << n
    k = len(targets)
    m = np.asarray(matrix, dtype=np.complex128)
    full = np.zeros((dim, dim), dtype=np.complex128)
    tmask = 0
    for t in targets:
        tmask |= 1 << t
    for col in range(dim):
        x = 0
        for i, t in enumerate(targets):
            x |= ((col >> t) & 1) << i
        base = col & ~tmask
        for y in range(1 << k):
            row = base
            for i, t in enumerate(targets):
                row |= ((y >> i) & 1) << t
            full[row, col] += m[y, x]
    return full


def controlled(n: int, controls, cvals, targets, matrix) -> np.ndarray:
    """Full operator of a controlled k-qubit matrix (identity off the control subspace)."""
    dim = 1 << n
    full = expand(n, targets, matrix)
    out = np.eye(dim, dtype=np.complex128)
    for col in range(dim):
        if all(((col >> c) & 1) == v for c, v in zip(controls, cvals)):
            out[:, col] = full[:, col]
    return out


def pauli_matrix(n: int, targets, paulis) -> np.ndarray:
    m = np.eye(1, dtype=np.complex128)
    ops = {t: p for t, p in zip(targets, paulis)}
    for q in reversed(range(n)):  # qubit n-1 is the leftmost Kronecker factor
        m = np.kron(m, PAULI[ops.get(q, 0)])
    return m


def host_hypot(x: float, y: float) -> float:
    """|x + iy| as the reference's std::abs(std::complex<double>) rounds it on
    glibc 2.35 (libstdc++ calls cabs = hypot): C. F. Borges, "An Improved
    Algorithm for hypot(a, b)", arXiv:1904.09481 (2019), non-FMA form, with a
    power-of-two rescale above 2^511 / below 2^-511 and the shortcut ax + ay when
    the operands differ by more than 2^54.  Python floats round every operation
    to nearest without contraction, as that code does."""
    import math

    ax, ay = abs(x), abs(y)
    if ax < ay:
        ax, ay = ay, ax
    if math.isinf(ax) or math.isinf(ay):
        return math.inf
    if math.isnan(ax) or math.isnan(ay):
        return math.nan
    eps, scale, large, tiny = 2.0 ** -54, 2.0 ** -600, 2.0 ** 511, 2.0 ** -511

    def kernel(a: float, b: float) -> float:
        h = math.sqrt(a * a + b * b)
        if h <= 2.0 * b:
            delta = h - b
            t1 = a * (2.0 * delta - a)
            t2 = (delta - 2.0 * (a - b)) * delta
        else:
            delta = h - a
            t1 = 2.0 * delta * (a - 2.0 * b)
            t2 = (4.0 * delta - b) * b + delta * delta
        return h - (t1 + t2) / (2.0 * h)

    if ax > large:
        if ay <= ax * eps:
            return ax + ay
        return kernel(ax * scale, ay * scale) / scale
    if ay < tiny:
        if ax >= ay / eps:
            return ax + ay
        return kernel(ax / scale, ay / scale) * scale
    if ay <= ax * eps:
        return ax + ay
    return kernel(ax, ay)
