"""QuantumCircuit.update_quantum_state through the stage executor: runs of fusable
gates (dense / diagonal on <= 4 qubits counting controls, CZ, CNOT) become compiled
programs, the rest run gate by gate (src/cppsim/circuit.cpp:46-57 is the per-gate
loop). Every check compares against the same circuit with fusion switched off, which
is one device call per gate."""
from __future__ import annotations

import numpy as np
import pytest

import qulacs_amd as q
from qulacs_amd import gate
from tests.golden.gen import haar_unitary

pytestmark = pytest.mark.gpu


def _build(n, seed, depth=6):
    rng = np.random.default_rng(seed)
    c = q.ParametricQuantumCircuit(n)
    for layer in range(depth):
        for i in range(n):
            c.add_gate(gate.DenseMatrix(i, haar_unitary(rng, 1)))
            c.add_parametric_RZ_gate(i, float(rng.uniform(0, 6.28)))
        for i in range(layer % 2, n - 1, 2):
            c.add_CNOT_gate(i, i + 1)
        a, b = rng.choice(n, 2, replace=False)
        c.add_CZ_gate(int(a), int(b))
        t = rng.choice(n, 3, replace=False).tolist()
        c.add_gate(gate.DenseMatrix(t, haar_unitary(rng, 3)))
        t0 = int(rng.integers(n))
        g = gate.to_matrix_gate(gate.RX(t0, 0.3))
        g.add_control_qubit((t0 + 1) % n, 1)  # controlled rotation: a 2-qubit dense gate
        c.add_gate(g)
        c.add_gate(gate.PauliRotation([0, n - 1], [1, 2], 0.7))
        c.add_gate(gate.DiagonalMatrix([1, 2], np.exp(1j * rng.uniform(0, 6, 4))))
        if layer == 2:
            c.add_gate(gate.DenseMatrix(rng.choice(n, 5, replace=False).tolist(), haar_unitary(rng, 5)))  # not fusable
            c.add_gate(gate.TOFFOLI(0, 1, 2))
            c.add_gate(gate.FusedSWAP(0, 2, 2))
    return c


def _run(c, n, fused):
    q.set_circuit_fusion(fused)
    try:
        s = q.QuantumState(n)
        s.set_zero_state()
        c.update_quantum_state(s)
        return s.get_vector()
    finally:
        q.set_circuit_fusion(True)


@pytest.mark.parametrize("n", [6, 12, 21, 22])
def test_fused_equals_per_gate(n):
    c = _build(n, n, depth=6 if n < 22 else 3)
    ref = _run(c, n, False)
    got = _run(c, n, True)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12
    got2 = _run(c, n, True)  # the compiled plan is reused
    np.testing.assert_array_equal(got2, got)


def test_plan_follows_parameters_and_edits():
    n = 10
    c = _build(n, 3)
    base = _run(c, n, True)
    c.set_parameter(0, 1.234)
    ref = _run(c, n, False)
    got = _run(c, n, True)
    assert np.linalg.norm(got - ref) < 1e-12 and np.linalg.norm(got - base) > 1e-6
    c.add_H_gate(4)
    c.remove_gate(3)
    np.testing.assert_allclose(_run(c, n, True), _run(c, n, False), atol=1e-12)
    q.circuit.QuantumCircuitOptimizer().optimize_light(c)
    np.testing.assert_allclose(_run(c, n, True), _run(c, n, False), atol=1e-12)


def test_noise_gates_split_the_runs():
    n = 8
    out = []
    for fused in (False, True):
        c = q.QuantumCircuit(n)
        for i in range(n):
            c.add_H_gate(i)
            c.add_RY_gate(i, 0.1 * i)
        dep = gate.DepolarizingNoise(3, 0.5)
        dep.set_seed(7)
        c.add_gate(dep)
        for i in range(n - 1):
            c.add_CNOT_gate(i, i + 1)
            c.add_RZ_gate(i, 0.2)
        out.append(_run(c, n, fused))
    np.testing.assert_allclose(out[1], out[0], atol=1e-12)


def test_density_matrix_circuit_matches_the_pure_state():
    n = 5
    c = _build(n, 9, depth=2)
    d = q.DensityMatrix(n)
    d.set_zero_state()
    c.update_quantum_state(d)
    v = _run(c, n, False)
    np.testing.assert_allclose(d.get_matrix(), np.outer(v, v.conj()), atol=1e-12)


def test_short_runs_from_22_qubits():
    # from 22 qubits a run of two fusable gates between unfusable ones is a program
    n = 22
    rng = np.random.default_rng(5)
    c = q.QuantumCircuit(n)
    for k in range(3):
        c.add_gate(gate.DenseMatrix(rng.choice(n, 5, replace=False).tolist(), haar_unitary(rng, 5)))
        c.add_H_gate(k)
        c.add_RZ_gate(20 - k, 0.3 + k)
    np.testing.assert_allclose(_run(c, n, True), _run(c, n, False), atol=1e-12)


def _dm_run(c, n, fused, v0):
    q.set_circuit_fusion(fused)
    try:
        d = q.DensityMatrix(n)
        d.load(v0)
        c.update_quantum_state(d)
        return d.get_matrix()
    finally:
        q.set_circuit_fusion(True)


def test_density_matrix_superoperator_fusion():
    # gates and channels on <= 3 qubits together compose into one superoperator pass
    n = 7
    rng = np.random.default_rng(21)
    c = q.QuantumCircuit(n)
    for layer in range(4):
        for i in range(n):
            c.add_gate(gate.DenseMatrix(i, haar_unitary(rng, 1)))
            c.add_gate(gate.DepolarizingNoise(i, 0.05))
        for i in range(layer % 2, n - 1, 2):
            c.add_CNOT_gate(i, i + 1)
            c.add_gate(gate.TwoQubitDepolarizingNoise(i, i + 1, 0.02))
        c.add_gate(gate.AmplitudeDampingNoise(3, 0.1))
        c.add_gate(gate.TOFFOLI(0, 1, 2))
        c.add_gate(gate.DenseMatrix([1, 4, 6, 2], haar_unitary(rng, 4)))  # 4 qubits: its own pass
        c.add_gate(gate.Pauli([0, 5], [2, 3]))
        c.add_gate(gate.DiagonalMatrix([2, 5], np.exp(1j * rng.uniform(0, 6, 4))))
    v0 = q.QuantumState(n)
    v0.set_Haar_random_state(4)
    v = v0.get_vector()
    ref = _dm_run(c, n, False, v)
    got = _dm_run(c, n, True, v)
    np.testing.assert_allclose(got, ref, atol=1e-12)
    assert abs(np.trace(got) - 1) < 1e-12
