"""Several single-qubit channels on a density matrix in one pass
(qm_dm_single_qubit_channels_host) against the Kraus sum computed with numpy,
and the density-matrix circuit loop that uses it for runs of three or more such
channels against the same circuit applied one gate at a time."""
from __future__ import annotations

import numpy as np
import pytest

import qulacs_amd as q
from qulacs_amd import _lib, gate
from tests.golden.gen import haar_unitary
from tests.gpu_util import DevState, cmat, u32

pytestmark = pytest.mark.gpu
TOL = 1e-12

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.diag([1.0, -1.0]).astype(complex)
I2 = np.eye(2, dtype=complex)


def _random_channel(rng):
    """Kraus operators of a random channel that maps {00, 11} and {01, 10} onto themselves."""
    kind = int(rng.integers(4))
    if kind == 0:  # Pauli channel
        p = rng.dirichlet(np.ones(4))
        return [np.sqrt(w) * m for w, m in zip(p, (I2, X, Y, Z))]
    if kind == 1:  # amplitude damping
        g = rng.uniform()
        return [np.array([[1, 0], [0, np.sqrt(1 - g)]], dtype=complex), np.array([[0, np.sqrt(g)], [0, 0]], dtype=complex)]
    if kind == 2:  # phase damping with a phase
        lam = rng.uniform()
        ph = np.exp(1j * rng.uniform(0, 2 * np.pi))
        return [np.diag([1, np.sqrt(1 - lam) * ph]), np.diag([0, np.sqrt(lam)]).astype(complex)]
    # generalized amplitude damping
    g, p = rng.uniform(size=2)
    return [np.sqrt(p) * np.array([[1, 0], [0, np.sqrt(1 - g)]]), np.sqrt(p) * np.array([[0, np.sqrt(g)], [0, 0]]),
            np.sqrt(1 - p) * np.array([[np.sqrt(1 - g), 0], [0, 1]]), np.sqrt(1 - p) * np.array([[0, 0], [np.sqrt(g), 0]])]


def _superop(kraus):
    """4 x 4 superoperator, member = row bit * 2 + column bit."""
    s = np.zeros((4, 4), dtype=complex)
    for k in kraus:
        for ro in range(2):
            for co in range(2):
                for ri in range(2):
                    for ci in range(2):
                        s[ro * 2 + co, ri * 2 + ci] += k[ro, ri] * np.conj(k[co, ci])
    return s


def _embed(k, qb, n):
    ops = [I2] * n
    ops[n - 1 - qb] = k
    out = ops[0]
    for o in ops[1:]:
        out = np.kron(out, o)
    return out


def _random_dm(rng, n):
    a = rng.standard_normal((1 << n, 1 << n)) + 1j * rng.standard_normal((1 << n, 1 << n))
    rho = a @ a.conj().T
    return rho / np.trace(rho)


@pytest.mark.parametrize("n", [3, 5, 8])
def test_channels_in_one_pass_match_kraus_sums(n):
    rng = np.random.default_rng(40 + n)
    dim = 1 << n
    for count in range(1, min(5, n) + 1):
        for _ in range(2):
            qubits = rng.choice(n, count, replace=False).tolist()
            chans = [_random_channel(rng) for _ in qubits]
            rho = _random_dm(rng, n)
            want = rho.copy()
            for qb, ks in zip(qubits, chans):
                want = sum(_embed(k, qb, n) @ want @ _embed(k, qb, n).conj().T for k in ks)
            print(f"{n} qubits, channels on {qubits}", flush=True)
            with DevState.from_host(rho.ravel()) as s:
                ops = np.concatenate([_superop(ks).ravel() for ks in chans])
                _lib.call("qm_dm_single_qubit_channels_host", u32(qubits), count, cmat(ops), s.ptr, dim, None, 0)
                got = s.download().reshape(dim, dim)
            np.testing.assert_allclose(got, want, atol=TOL)


def test_channels_reject_mixing_superoperators():
    rng = np.random.default_rng(5)
    n = 4
    dim = 1 << n
    h = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2)
    with DevState.from_host(_random_dm(rng, n).ravel()) as s:
        with pytest.raises(Exception, match="mixes"):
            _lib.call("qm_dm_single_qubit_channels_host", u32([1]), 1, cmat(_superop([h]).ravel()), s.ptr, dim, None, 0)
        with pytest.raises(Exception, match="listed twice"):
            ops = np.concatenate([_superop([I2]).ravel()] * 2)
            _lib.call("qm_dm_single_qubit_channels_host", u32([2, 2]), 2, cmat(ops), s.ptr, dim, None, 0)


@pytest.mark.parametrize("n", [6, 9])
def test_noisy_circuit_loop_matches_gate_by_gate(n):
    """5-qubit gates followed by depolarizing, amplitude damping and dephasing on
    each of their qubits: the circuit loop (channel runs in one pass, pairs fused
    into superoperators) equals the circuit applied one gate at a time."""
    rng = np.random.default_rng(n)
    circ = q.QuantumCircuit(n)
    for layer in range(4):
        tg = rng.choice(n, 5, replace=False).tolist()
        circ.add_gate(gate.DenseMatrix(tg, haar_unitary(rng, 5)))
        for j, t in enumerate(tg):
            p = float(rng.uniform(0.0, 0.2))
            if (layer + j) % 3 == 0:
                circ.add_gate(gate.DepolarizingNoise(t, p))
            elif (layer + j) % 3 == 1:
                circ.add_gate(gate.AmplitudeDampingNoise(t, p))
            else:
                circ.add_gate(gate.DephasingNoise(t, p))
        a = int(rng.integers(n))
        circ.add_gate(gate.DenseMatrix([a], haar_unitary(rng, 1)))
        circ.add_gate(gate.DepolarizingNoise(a, 0.05))
    fused = q.DensityMatrix(n)
    fused.set_Haar_random_state(3)
    plain = fused.copy()
    circ.update_quantum_state(fused)
    q.set_circuit_fusion(False)
    try:
        for i in range(circ.get_gate_count()):
            circ.get_gate(i).update_quantum_state(plain)
    finally:
        q.set_circuit_fusion(True)
    np.testing.assert_allclose(fused.get_matrix(), plain.get_matrix(), atol=TOL)
    assert abs(np.trace(fused.get_matrix()).real - 1) < TOL
