"""Channels on density matrices, as test/cppsim/test_noise_dm.cpp checks them for the
reference: Probabilistic and CPTP maps over random unitaries against
sum_i p_i U rho U^dagger; amplitude damping from explicit Kraus matrices; the
depolarizing closed forms (DepolarizingNoise(3p/4) and TwoQubitDepolarizingNoise(15p/16)
are (1-p) rho + p I/d); Instrument and Adaptive on a density matrix; samples of a
one-qubit density matrix lie in {0, 1}.  Every result is also trace preserving."""
from __future__ import annotations

import numpy as np
import pytest

import qulacs_amd as q
from qulacs_amd import gate

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _haar_dm(n, seed):
    dm = q.DensityMatrix(n)
    dm.set_Haar_random_state(seed)
    return dm


def _trace(dm):
    return np.trace(dm.get_matrix()).real


@pytest.mark.parametrize("repeat", range(4))
def test_probabilistic_mixture_of_random_unitaries(repeat):
    n, count = 5, 5
    rng = np.random.default_rng(repeat)
    probs = rng.uniform(size=count)
    probs /= probs.sum()
    us = [gate.RandomUnitary(list(range(n)), 100 * repeat + i) for i in range(count)]
    dm = _haar_dm(n, repeat)
    rho = dm.get_matrix()
    want = sum(p * u.get_matrix() @ rho @ u.get_matrix().conj().T for p, u in zip(probs, us))
    gate.Probabilistic(probs.tolist(), us).update_quantum_state(dm)
    np.testing.assert_allclose(dm.get_matrix(), want, atol=TOL)
    assert abs(_trace(dm) - 1) < TOL


@pytest.mark.parametrize("repeat", range(4))
def test_cptp_of_scaled_random_unitaries(repeat):
    n, count = 2, 5
    rng = np.random.default_rng(repeat + 10)
    probs = rng.uniform(size=count)
    probs /= probs.sum()
    kraus = []
    for i, p in enumerate(probs):
        k = gate.RandomUnitary(list(range(n)), 50 * repeat + i)
        k.multiply_scalar(np.sqrt(p))
        kraus.append(k)
    dm = _haar_dm(n, repeat + 20)
    rho = dm.get_matrix()
    want = sum(k.get_matrix() @ rho @ k.get_matrix().conj().T for k in kraus)
    gate.CPTP(kraus).update_quantum_state(dm)
    np.testing.assert_allclose(dm.get_matrix(), want, atol=TOL)
    assert abs(_trace(dm) - 1) < TOL


@pytest.mark.parametrize("prob", [0.0, 0.3, 0.85, 1.0])
def test_amplitude_damping_from_kraus_matrices(prob):
    k0 = np.array([[1, 0], [0, np.sqrt(1 - prob)]])
    k1 = np.array([[0, np.sqrt(prob)], [0, 0]])
    dm = _haar_dm(1, 7)
    rho = dm.get_matrix()
    want = k0 @ rho @ k0.conj().T + k1 @ rho @ k1.conj().T
    gate.CPTP([gate.DenseMatrix(0, k0), gate.DenseMatrix(0, k1)]).update_quantum_state(dm)
    np.testing.assert_allclose(dm.get_matrix(), want, atol=TOL)
    dm2 = _haar_dm(1, 7)
    gate.AmplitudeDampingNoise(0, prob).update_quantum_state(dm2)
    np.testing.assert_allclose(dm2.get_matrix(), want, atol=TOL)


@pytest.mark.parametrize("prob", [0.2, 0.6])
def test_depolarizing_closed_forms(prob):
    for n, make in ((1, lambda: gate.DepolarizingNoise(0, prob * 3 / 4)),
                    (2, lambda: gate.TwoQubitDepolarizingNoise(0, 1, prob * 15 / 16))):
        dim = 1 << n
        dm = _haar_dm(n, n)
        rho = dm.get_matrix()
        make().update_quantum_state(dm)
        np.testing.assert_allclose(dm.get_matrix(), (1 - prob) * rho + prob / dim * np.eye(dim), atol=TOL)
        assert abs(_trace(dm) - 1) < TOL


def test_instrument_on_density_matrix_applies_the_whole_map():
    # four projectors onto |q1 q0> basis states, in the reference's order
    p0, p1 = gate.P0, gate.P1
    kraus = [gate.merge(p1(0), p0(1)), gate.merge(p0(0), p1(1)), gate.merge(p0(0), p0(1)),
             gate.merge(p1(0), p1(1))]
    dm = _haar_dm(3, 5)
    rho = dm.get_matrix()
    gate.Instrument(kraus, 1).update_quantum_state(dm)
    # dephasing of qubits 0 and 1: blocks with equal (q1, q0) on both sides survive
    idx = np.arange(8) & 3
    np.testing.assert_allclose(dm.get_matrix(), np.where(idx[:, None] == idx[None, :], rho, 0), atol=TOL)
    assert abs(_trace(dm) - 1) < TOL


def test_adaptive_on_density_matrix():
    x = gate.X(0)
    adaptive = gate.Adaptive(x, lambda reg: reg[2] == 1)
    dm = q.DensityMatrix(1)
    dm.set_computational_basis(0)
    dm.set_classical_value(2, 1)
    adaptive.update_quantum_state(dm)
    np.testing.assert_allclose(dm.get_matrix(), [[0, 0], [0, 1]], atol=0)
    dm.set_classical_value(2, 0)
    adaptive.update_quantum_state(dm)
    np.testing.assert_allclose(dm.get_matrix(), [[0, 0], [0, 1]], atol=0)


def test_density_matrix_samples_are_basis_indices():
    dm = _haar_dm(1, 9)
    samples = dm.sampling(1000, 3)
    assert len(samples) == 1000 and set(samples) <= {0, 1}
    p1 = dm.get_matrix()[1, 1].real
    assert abs(np.mean(samples) - p1) < 0.06
