"""Quantum-trajectory evolution (src/cppsim/gate_noisy_evolution.cpp) on the device:
without collapse operators both gates are exp(-i H t) (RK4 within its step error,
the diagonalised form to rounding), checked against scipy's expm; with amplitude
damping sqrt(g) |0><1| the trajectory average of the excited population follows
exp(-g t), as the Lindblad equation gives (a statistical check, 4 sigma)."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

import qulacs_amd as q
from qulacs_amd import gate

pytestmark = pytest.mark.gpu


def _hamiltonian(n):
    h = q.Observable(n)
    h.add_operator(0.7, "X 0 X 1")
    h.add_operator(-0.4, "Z 1")
    h.add_operator(0.25, "Y 0 Z 2")
    return h


@pytest.mark.parametrize("fast", [False, True])
def test_no_collapse_is_unitary_evolution(fast):
    n, t = 3, 0.8
    h = _hamiltonian(n)
    s = q.QuantumState(n)
    s.set_Haar_random_state(4)
    v = s.get_vector()
    g = gate.NoisyEvolution_fast(h, [], t) if fast else gate.NoisyEvolution(h, [], t, 0.01)
    g.update_quantum_state(s)
    want = scipy.linalg.expm(-1j * t * h.get_matrix().toarray()) @ v
    np.testing.assert_allclose(s.get_vector(), want, atol=1e-10 if fast else 1e-7)


def _decay_op(gamma):
    c = q.GeneralQuantumOperator(1)
    c.add_operator(0.5 * np.sqrt(gamma), "X 0")
    c.add_operator(0.5j * np.sqrt(gamma), "Y 0")
    return c


@pytest.mark.parametrize("fast", [False, True])
def test_amplitude_damping_trajectories(fast):
    gamma, t, shots = 1.0, 1.0, 200
    h = q.Observable(1)
    h.add_operator(0.3, "Z 0")
    c = _decay_op(gamma)
    g = gate.NoisyEvolution_fast(h, [c], t) if fast else gate.NoisyEvolution(h, [c], t, 0.02)
    g.set_seed(123)
    excited = 0
    for _ in range(shots):
        s = q.QuantumState(1)
        s.set_computational_basis(1)
        g.update_quantum_state(s)
        v = s.get_vector()
        assert abs(np.vdot(v, v) - 1) < 1e-9
        p1 = abs(v[1]) ** 2
        assert p1 < 1e-9 or p1 > 1 - 1e-9  # a trajectory is either decayed or not
        excited += p1 > 0.5
    p = np.exp(-gamma * t)
    assert abs(excited / shots - p) < 4 * np.sqrt(p * (1 - p) / shots)


def test_effective_hamiltonian_and_json():
    h = q.Observable(1)
    h.add_operator(0.3, "Z 0")
    c = _decay_op(0.5)
    for g in (gate.NoisyEvolution(h, [c], 1.0, 0.01), gate.NoisyEvolution_fast(h, [c], 1.0)):
        heff = g.get_effective_hamiltonian().get_matrix().toarray()
        sm = np.array([[0, np.sqrt(0.5)], [0, 0]])
        want = 0.3 * np.diag([1, -1]) - 0.5j * sm.conj().T @ sm
        np.testing.assert_allclose(heff, want, atol=1e-15)
        g2 = gate.from_json(g.to_json())
        assert g2.get_name() == g.get_name()
        np.testing.assert_allclose(g2.get_effective_hamiltonian().get_matrix().toarray(), want, atol=1e-15)
