"""Operator actions and ground-state solvers on the device, as
test/cppsim/test_hamiltonian.cpp:395-560 checks them for the reference: the
eigenvalue each solver returns against numpy's eigvalsh of get_matrix(), and the
state it leaves behind against the eigen-equation A|q> = lambda|q> (both within
the reference's tolerances: 1e-2 for the power method, 1e-6 for Arnoldi and
Lanczos); apply_to_state (both forms) against the matrix-vector product; the H2
Hamiltonian read from OpenFermion text against the expectation value of its
matrix."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import qulacs_amd as q
from qulacs_amd import observable

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / "golden"


def _random_observable(n, count, seed, identity=None):
    o = q.Observable(n)
    o.add_random_operator(count, seed)
    if identity is not None:
        o.add_operator(q.PauliOperator(list(range(n)), [0] * n, identity))
    return o


def _check_eigenpair(o, st, value, eps):
    m = o.get_matrix().toarray()
    ground = np.linalg.eigvalsh(m)[0]
    assert abs(complex(value).real - ground) < eps, (value, ground)
    v = st.get_vector()
    av = m @ v
    lv = complex(value) * v
    av /= np.linalg.norm(av)
    lv /= np.linalg.norm(lv)
    assert np.max(np.abs(av - lv)) < eps


@pytest.mark.parametrize("n", [3, 4, 5])
def test_apply_to_state_both_forms(n):
    rng = np.random.default_rng(n)
    o = q.GeneralQuantumOperator(n)
    for _ in range(5):
        o.add_operator(q.PauliOperator(list(range(n)), rng.integers(0, 4, size=n).tolist(),
                                       complex(rng.normal(), rng.normal())))
    src = q.QuantumState(n)
    src.set_Haar_random_state(7)
    v = src.get_vector()
    want = o.get_matrix().toarray() @ v
    dst = q.QuantumState(n)
    work = q.QuantumState(n)
    o.apply_to_state(work, src, dst)
    np.testing.assert_allclose(dst.get_vector(), want, atol=1e-12)
    dst2 = q.QuantumState(n)
    o.apply_to_state(src, dst2)
    np.testing.assert_allclose(dst2.get_vector(), want, atol=1e-12)
    np.testing.assert_allclose(src.get_vector(), v, atol=0)  # restored: each Pauli undone


@pytest.mark.parametrize("seed", range(4))
def test_power_method(seed):
    rng = np.random.default_rng(seed)
    o = _random_observable(4, int(rng.integers(2, 12)), seed)
    st = q.QuantumState(4)
    st.set_Haar_random_state(seed + 100)
    value = o.solve_ground_state_eigenvalue_by_power_method(st, 500)
    # the power method converges to the lowest eigenvector only when the spectrum's
    # gap allows within 500 steps; the reference accepts one miss in ten
    m = o.get_matrix().toarray()
    w = np.linalg.eigvalsh(m)
    if len(w) > 1 and abs(w[1] - w[0]) < 1e-3:
        pytest.skip("degenerate ground space")
    _check_eigenpair(o, st, value, 1e-2)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("identity", [None, 0.37])
def test_arnoldi_method(seed, identity):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 7))
    o = _random_observable(n, int(rng.integers(2, 12)), seed + 10, identity)
    st = q.QuantumState(n)
    st.set_Haar_random_state(seed + 200)
    iters = min(60, 1 << n)
    value = o.solve_ground_state_eigenvalue_by_arnoldi_method(st, iters)
    _check_eigenpair(o, st, value, 1e-6)


@pytest.mark.parametrize("seed", range(6))
def test_lanczos_method(seed):
    rng = np.random.default_rng(seed + 50)
    n = int(rng.integers(3, 7))
    o = _random_observable(n, int(rng.integers(2, 12)), seed + 20)
    st = q.QuantumState(n)
    st.set_Haar_random_state(seed + 300)
    # Lanczos without re-orthogonalisation: a Krylov space no larger than the state
    value = o.solve_ground_state_eigenvalue_by_lanczos_method(st, min(70, 1 << n))
    assert abs(complex(value).real - np.linalg.eigvalsh(o.get_matrix().toarray())[0]) < 1e-6


def test_solvers_with_explicit_shift():
    o = _random_observable(4, 6, 5)
    ground = np.linalg.eigvalsh(o.get_matrix().toarray())[0]
    st = q.QuantumState(4)
    st.set_Haar_random_state(1)
    assert abs(o.solve_ground_state_eigenvalue_by_arnoldi_method(st, 16, 3.0).real - ground) < 1e-6
    st.set_Haar_random_state(1)
    assert abs(o.solve_ground_state_eigenvalue_by_lanczos_method(st, 16, 3.0).real - ground) < 1e-6
    with pytest.raises(RuntimeError):
        q.Observable(2).solve_ground_state_eigenvalue_by_power_method(q.QuantumState(2), 3)


def test_h2_expectation_from_openfermion_text():
    o = observable.create_observable_from_openfermion_text((GOLDEN / "openfermion_H2.txt").read_text())
    m = o.get_matrix().toarray()
    st = q.QuantumState(4)
    st.set_computational_basis(0)
    assert abs(o.get_expectation_value(st) - m[0, 0].real) < 1e-12
    st.set_Haar_random_state(3)
    v = st.get_vector()
    assert abs(o.get_expectation_value(st) - np.vdot(v, m @ v).real) < 1e-12
    # H2's ground energy (STO-3G, 0.74 A) from the Lanczos solver
    st.set_Haar_random_state(4)
    e = o.solve_ground_state_eigenvalue_by_lanczos_method(st, 16).real
    assert abs(e - np.linalg.eigvalsh(m)[0]) < 1e-8


def test_pauli_beyond_the_state_is_an_error():
    # test/cppsim/test_pauli_operator.cpp:20-30
    st = q.QuantumState(2)
    st.set_Haar_random_state(1)
    with pytest.raises(Exception):
        q.PauliOperator("X 0 X 1 X 3", 2.0).get_expectation_value(st)
    assert abs(q.PauliOperator("X 0 X 1", 2.0).get_expectation_value(st)) <= 2.0 + 1e-12


def test_observable_on_a_larger_state_and_identity_term():
    # test/cppsim/test_hamiltonian.cpp:566-639: an observable read from OpenFermion
    # text acts on a state with two more qubits as the sum of its terms; an "I 0"
    # term applied to a state scales it by its coefficient
    o = observable.create_observable_from_openfermion_text((GOLDEN / "openfermion_H2.txt").read_text())
    st = q.QuantumState(o.get_qubit_count() + 2)
    for seed in (None, 5):
        if seed is None:
            st.set_computational_basis(0)
        else:
            st.set_Haar_random_state(seed)
        want = sum(o.get_term(i).get_expectation_value(st) for i in range(o.get_term_count()))
        got = o.get_expectation_value(st)
        assert abs(got - want) < 1e-12 and abs(complex(want).imag) < 1e-12
    ident = q.Observable(3)
    ident.add_operator(0.5, "I 0")
    src, dst = q.QuantumState(3), q.QuantumState(3)
    src.set_Haar_random_state(2)
    ident.apply_to_state(src, dst)
    src.add_state_with_coef(-1 / 0.5, dst)
    assert src.get_squared_norm() < 1e-24
