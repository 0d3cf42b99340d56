"""State-level closed forms from test/cppsim/test_state.cpp and test_state_dm.cpp:
zero probabilities of sqrt(i)-weighted basis states (30/55, 27/55, 33/55, 28/55),
the marginal-probability table of a two-qubit state (0, 1 and 2 = "either"),
set_zero_norm_state, add_state(_with_coef), multiply_coef and make_superposition
against numpy, on state vectors and density matrices."""
from __future__ import annotations

import numpy as np
import pytest

import qulacs_amd as q
from qulacs_amd import state as S

pytestmark = pytest.mark.gpu
EPS = 1e-12


def test_zero_probability_closed_form():
    n = 10
    st = q.QuantumState(n)
    st.set_computational_basis(1)
    for i in range(2, 11):
        tmp = q.QuantumState(n)
        tmp.set_computational_basis(i)
        st.add_state_with_coef(np.sqrt(i), tmp)
    st.normalize(st.get_squared_norm())
    for qubit, num in enumerate((30, 27, 33, 28)):
        assert abs(st.get_zero_probability(qubit) - num / 55) < EPS


@pytest.mark.parametrize("dm", [False, True])
def test_marginal_probability_table(dm):
    sv = q.QuantumState(2)
    sv.set_Haar_random_state(5)
    p = np.abs(sv.get_vector()) ** 2
    st = sv
    if dm:
        st = q.DensityMatrix(2)
        st.load(sv)
    table = {(0, 0): p[0], (1, 0): p[1], (0, 1): p[2], (1, 1): p[3], (0, 2): p[0] + p[2],
             (1, 2): p[1] + p[3], (2, 0): p[0] + p[1], (2, 1): p[2] + p[3], (2, 2): 1.0}
    for key, want in table.items():
        assert abs(st.get_marginal_probability(list(key)) - want) < EPS, key


def test_zero_norm_state_and_arithmetic():
    n = 5
    st = q.QuantumState(n)
    st.set_Haar_random_state(1)
    st.set_zero_norm_state()
    assert not np.any(st.get_vector())
    a, b = q.QuantumState(n), q.QuantumState(n)
    a.set_Haar_random_state(2)
    b.set_Haar_random_state(3)
    va, vb = a.get_vector(), b.get_vector()
    a.add_state(b)
    np.testing.assert_allclose(a.get_vector(), va + vb, atol=EPS)
    a.add_state_with_coef(0.5 - 0.25j, b)
    np.testing.assert_allclose(a.get_vector(), va + vb + (0.5 - 0.25j) * vb, atol=EPS)
    a.multiply_coef(0.3 + 0.7j)
    np.testing.assert_allclose(a.get_vector(), (0.3 + 0.7j) * (va + vb + (0.5 - 0.25j) * vb), atol=EPS)


def test_make_superposition_and_mixture():
    n = 4
    rng = np.random.default_rng(4)
    s1, s2 = q.QuantumState(n), q.QuantumState(n)
    s1.set_Haar_random_state(6)
    s2.set_Haar_random_state(7)
    c1, c2 = complex(*rng.uniform(size=2)), complex(*rng.uniform(size=2))
    sup = S.make_superposition(c1, s1, c2, s2)
    np.testing.assert_allclose(sup.get_vector(), c1 * s1.get_vector() + c2 * s2.get_vector(), atol=EPS)
    d1, d2 = q.DensityMatrix(n), q.DensityMatrix(n)
    d1.set_Haar_random_state(8)
    d2.set_Haar_random_state(9)
    mix = S.make_mixture(c1, d1, c2, d2)
    np.testing.assert_allclose(mix.get_matrix(), c1 * d1.get_matrix() + c2 * d2.get_matrix(), atol=EPS)


def test_density_matrix_add_and_multiply():
    n = 3
    a, b = q.DensityMatrix(n), q.DensityMatrix(n)
    a.set_Haar_random_state(10)
    b.set_Haar_random_state(11)
    ma, mb = a.get_matrix(), b.get_matrix()
    a.add_state(b)
    np.testing.assert_allclose(a.get_matrix(), ma + mb, atol=EPS)
    a.multiply_coef(0.25)
    np.testing.assert_allclose(a.get_matrix(), 0.25 * (ma + mb), atol=EPS)
