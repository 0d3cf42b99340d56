"""JSON wire format of device states (python/tests/single_cpu/test_json.py of the
reference, on QuantumStateMI355 / DensityMatrixMI355): exact round trips of the
amplitudes and the classical register, pickling, and gates / circuits restored
from JSON acting identically on Haar-random states."""
from __future__ import annotations

import json
import pickle

import numpy as np
import pytest

import qulacs_amd as q
from qulacs_amd import QuantumCircuit, gate, state

pytestmark = pytest.mark.gpu


def test_state_vector_round_trip_is_exact():
    s = q.QuantumState(5)
    s.set_Haar_random_state(7)
    s.set_classical_value(2, 1)
    d = json.loads(s.to_json())
    assert d["name"] == "QuantumState" and d["qubit_count"] == "5" and len(d["state_vector"]) == 32
    assert d["classical_register"] == ["0", "0", "1"]
    for s2 in (state.from_json(s.to_json()), pickle.loads(pickle.dumps(s))):
        assert isinstance(s2, q.QuantumStateMI355)
        assert np.array_equal(s2.get_vector(), s.get_vector())
        assert s2.get_classical_value(2) == 1


def test_density_matrix_round_trip_is_exact():
    r = q.DensityMatrix(3)
    r.set_Haar_random_state(3)
    gate.DepolarizingNoise(0, 0.2).update_quantum_state(r)
    d = json.loads(r.to_json())
    assert d["name"] == "DensityMatrix" and len(d["density_matrix"]) == 64
    for r2 in (state.from_json(r.to_json()), pickle.loads(pickle.dumps(r))):
        assert isinstance(r2, q.DensityMatrixMI355)
        assert np.array_equal(r2.get_matrix(), r.get_matrix())


def test_gates_from_json_act_identically():
    rng = np.random.default_rng(21)
    n = 3
    gates = [gate.Identity(0), gate.X(0), gate.Y(1), gate.H(2), gate.S(0), gate.Tdag(1), gate.sqrtY(2),
             gate.CNOT(0, 1), gate.CZ(1, 2), gate.SWAP(0, 2), gate.Pauli([0, 1], [1, 2]),
             gate.PauliRotation([0, 1], [1, 2], rng.uniform()), gate.merge(gate.X(0), gate.Y(1)),
             gate.P0(0), gate.P1(1), gate.U1(0, rng.uniform()), gate.U2(0, rng.uniform(), rng.uniform()),
             gate.U3(0, rng.uniform(), rng.uniform(), rng.uniform()), gate.RX(0, rng.uniform()),
             gate.RY(1, rng.uniform()), gate.RZ(2, rng.uniform()), gate.ParametricRX(0, rng.uniform()),
             gate.ParametricPauliRotation([0, 1], [1, 1], rng.uniform()), gate.RandomUnitary([0, 1, 2])]
    for g in gates:
        a = q.QuantumState(n)
        a.set_Haar_random_state(int(rng.integers(1 << 30)))
        b = a.copy()
        g.update_quantum_state(a)
        gate.from_json(g.to_json()).update_quantum_state(b)
        assert np.array_equal(a.get_vector(), b.get_vector()), g.get_name()


def test_circuit_from_json_acts_identically():
    n = 3
    c = QuantumCircuit(n)
    for i in range(3):
        c.add_gate(gate.RandomUnitary([0, 1, 2], i))
    a = q.QuantumState(n)
    a.set_Haar_random_state(5)
    b = a.copy()
    c.update_quantum_state(a)
    q.circuit.from_json(c.to_json()).update_quantum_state(b)
    assert np.array_equal(a.get_vector(), b.get_vector())


def test_measurement_writes_the_classical_register():
    """Measurement(q, reg) = Instrument([P0(q), P1(q)], reg) (src/cppsim/gate_factory.cpp:522-537):
    the outcome index goes to the register and the state collapses."""
    s = q.QuantumState(3)
    s.set_computational_basis(0b010)
    gate.Measurement(1, 0).update_quantum_state(s)
    gate.Measurement(0, 2).update_quantum_state(s)
    assert s.get_classical_value(0) == 1 and s.get_classical_value(2) == 0
    s = q.QuantumState(2)
    gate.H(0).update_quantum_state(s)
    counts = [0, 0]
    for seed in range(40):
        t = s.copy()
        gate.Measurement(0, 0, seed).update_quantum_state(t)
        k = t.get_classical_value(0)
        counts[k] += 1
        assert abs(t.get_zero_probability(0) - (1.0 - k)) < 1e-12  # collapsed onto the outcome
    assert counts[0] > 5 and counts[1] > 5
