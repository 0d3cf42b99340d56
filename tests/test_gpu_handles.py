"""Ownership of gates, states and operators across the Python boundary, as
python/tests/single_cpu/test_pointer_handling.py and test_utils.py check it for the
reference: a circuit copies what add_gate receives, so dropping the Python object
(and collecting it) leaves the circuit whole; get_gate returns an independent copy;
StateReflection keeps its own copy of the reflection state; a parametric circuit's
parameter positions follow insertions and removals; to_general_quantum_operator of a
random unitary acts like the gate.  Every circuit built this way is applied to a
device state and, where the gates are deterministic, checked against numpy."""
from __future__ import annotations

import gc

import numpy as np
import pytest
from scipy.sparse import lil_matrix

import qulacs_amd as q
from qulacs_amd import gate as G
from qulacs_amd import state as S

pytestmark = pytest.mark.gpu


def _every_gate_kind(ref_state):
    sparse = lil_matrix((4, 4), dtype=np.complex128)
    sparse[0, 0] = 1
    sparse[1, 1] = 1

    def shift(v, d):
        return (v + 1) % d

    gates = [
        G.Identity(0), G.X(0), G.Y(0), G.Z(0), G.H(0), G.S(0), G.Sdag(0), G.T(0), G.Tdag(0),
        G.sqrtX(0), G.sqrtXdag(0), G.sqrtY(0), G.sqrtYdag(0),
        G.Probabilistic([0.5, 0.5], [G.X(0), G.Y(0)]),
        G.CPTP([G.P0(0), G.P1(0)]),
        G.Instrument([G.P0(0), G.P1(0)], 1),
        G.Adaptive(G.X(0), lambda reg: True),
        G.CNOT(0, 1), G.CZ(0, 1), G.SWAP(0, 1), G.TOFFOLI(0, 1, 2), G.FREDKIN(0, 1, 2),
        G.Pauli([0, 1], [1, 2]), G.PauliRotation([0, 1], [1, 2], 0.1),
        G.DenseMatrix(0, np.eye(2)), G.DenseMatrix([0, 1], np.eye(4)),
        G.SparseMatrix([0, 1], sparse), G.DiagonalMatrix([0, 1], np.ones(4)),
        G.RandomUnitary([0, 1]), G.ReversibleBoolean([0, 1], shift), G.StateReflection(ref_state),
        G.BitFlipNoise(0, 0.1), G.DephasingNoise(0, 0.1), G.IndependentXZNoise(0, 0.1),
        G.DepolarizingNoise(0, 0.1), G.TwoQubitDepolarizingNoise(0, 1, 0.1),
        G.AmplitudeDampingNoise(0, 0.1), G.Measurement(0, 1),
        G.merge(G.X(0), G.Y(1)), G.add(G.X(0), G.Y(1)), G.to_matrix_gate(G.X(0)),
        G.P0(0), G.P1(0), G.U1(0, 0.0), G.U2(0, 0.0, 0.0), G.U3(0, 0.0, 0.0, 0.0),
        G.RX(0, 0.0), G.RY(0, 0.0), G.RZ(0, 0.0),
    ]
    gates.append(G.merge(gates[0], gates[1]))
    gates.append(G.add(gates[0], gates[1]))
    return gates


@pytest.mark.parametrize("parametric", [False, True])
def test_circuit_owns_added_gates(parametric):
    qc = q.ParametricQuantumCircuit(3) if parametric else q.QuantumCircuit(3)
    ref = q.QuantumState(3)
    gates = _every_gate_kind(ref)
    if parametric:
        gates += [G.ParametricRX(0, 0.1), G.ParametricRY(1, 0.2), G.ParametricRZ(2, 0.3),
                  G.ParametricPauliRotation([0, 1], [1, 1], 0.4)]
    del ref
    for g in gates + gates:
        if parametric and isinstance(g, q.QuantumGate_SingleParameter):
            qc.add_parametric_gate(g)
        else:
            qc.add_gate(g)
    count = len(gates) * 2
    del gates, g
    gc.collect()
    assert qc.get_gate_count() == count
    if parametric:
        assert qc.get_parameter_count() == 8
    st = q.QuantumState(3)
    qc.update_quantum_state(st)
    assert len(qc.to_string()) > 0
    del qc
    gc.collect()
    st.get_vector()  # the state outlives the circuit that acted on it


def test_get_gate_copies_outlive_their_circuit():
    def build():
        src = q.QuantumCircuit(2)
        src.add_X_gate(0)
        src.add_Y_gate(1)
        src.add_gate(G.RandomUnitary([0, 1], 5))
        out = q.QuantumCircuit(2)
        for i in range(src.get_gate_count()):
            out.add_gate(src.get_gate(i))
        want = q.QuantumState(2)
        src.update_quantum_state(want)
        return out, want.get_vector()

    out, want = build()
    gc.collect()
    st = q.QuantumState(2)
    out.update_quantum_state(st)
    np.testing.assert_allclose(st.get_vector(), want, atol=1e-14)
    # a gate handed out by get_gate is a copy: changing it leaves the circuit alone
    g = out.get_gate(0)
    assert g.get_name() == "X"
    out.remove_gate(0)
    assert g.get_target_index_list() == [0] and out.get_gate_count() == 2


def test_same_gate_added_many_times_and_noise_copies():
    st = q.QuantumState(1)
    circuit = q.QuantumCircuit(1)
    noise = G.DepolarizingNoise(0, 0)
    circuit.add_gate(noise)
    circuit.add_gate(noise.copy())
    circuit.add_gate(G.DephasingNoise(0, 0))
    circuit.add_gate(G.Probabilistic([0.1], [G.RX(0, 0)]))
    g = G.RX(0, 0.25)
    for _ in range(5):
        circuit.add_gate(g)
    del g, noise
    gc.collect()
    for _ in range(7):
        circuit.update_quantum_state(st)
    # zero-probability noise changes nothing, so 35 RX(0.25) rotations remain;
    # the Probabilistic gate's RX(0) is the identity whichever branch is drawn
    theta = 35 * 0.25
    np.testing.assert_allclose(st.get_vector(), [np.cos(theta / 2), 1j * np.sin(theta / 2)], atol=1e-12)
    assert "# of gate : 9" in circuit.to_string()


def test_observable_term_and_plain_gate_lifetimes():
    obs = q.Observable(1)
    obs.add_operator(1.0, "X 0")
    term = obs.get_term(0)
    del obs
    gc.collect()
    assert term.get_pauli_id_list() == [1] and term.get_coef() == 1.0
    for cls in (q.QuantumCircuit, q.ParametricQuantumCircuit):
        c = cls(1)
        x = G.X(0)
        c.add_gate(x)
        del x
        gc.collect()
        assert len(c.to_string()) > 0
        st = q.QuantumState(1)
        c.update_quantum_state(st)
        np.testing.assert_allclose(st.get_vector(), [0, 1], atol=0)


def test_state_reflection_keeps_its_own_state():
    n = 5

    def make():
        s2 = q.QuantumState(n)
        s2.set_Haar_random_state(3)
        v = s2.get_vector()
        gate = G.StateReflection(s2)
        del s2
        return gate, v

    gate, v = make()
    gc.collect()
    s1 = q.QuantumState(n)
    s1.set_Haar_random_state(4)
    psi = s1.get_vector()
    gate.update_quantum_state(s1)
    # 2|v><v| - I
    want = 2 * v * np.vdot(v, psi) - psi
    np.testing.assert_allclose(s1.get_vector(), want, atol=1e-12)


def test_sparse_matrix_from_lil():
    n = 5
    st = q.QuantumState(n)
    st.set_Haar_random_state(9)
    psi = st.get_vector()
    m = lil_matrix((4, 4), dtype=np.complex128)
    m[0, 0] = 1 + 1.0j
    m[1, 1] = 1.0 + 1.0j
    gate = G.SparseMatrix([0, 1], m)
    del m
    gate.update_quantum_state(st)
    want = psi.reshape(-1, 4).copy()
    want[:, :2] *= 1 + 1j
    want[:, 2:] = 0
    np.testing.assert_allclose(st.get_vector(), want.reshape(-1), atol=1e-12)


def test_copied_parametric_gate_values():
    def build():
        circuit = q.ParametricQuantumCircuit(1)
        g = G.ParametricRX(0, 0.1)
        circuit.add_parametric_gate(g)
        circuit.add_parametric_gate(g)
        circuit.add_gate(g)
        g.set_parameter_value(0.2)
        circuit.add_parametric_gate(g)
        circuit.add_parametric_RX_gate(0, 0.3)
        g2 = g.copy()
        g2.set_parameter_value(0.4)
        g.set_parameter_value(1.0)
        del g
        circuit.add_parametric_gate(g2)
        circuit.remove_gate(1)
        del g2
        return circuit

    c = build()
    gc.collect()
    for i in range(c.get_parameter_count()):
        v = c.get_parameter(i)
        assert v == pytest.approx(0.1 * i + 0.1)
        c.set_parameter(i, v + 1.0)
    d = c.copy()
    del c
    gc.collect()
    total = 0.1  # the unparametrised RX(0.1) added by add_gate
    for i in range(d.get_parameter_count()):
        d.set_parameter(i, d.get_parameter(i) + 10)
        assert d.get_parameter(i) == pytest.approx(11.1 + 0.1 * i)
        total += d.get_parameter(i)
    st = q.QuantumState(1)
    d.update_quantum_state(st)
    # every gate is an X rotation, so the angles add up
    np.testing.assert_allclose(st.get_vector(), [np.cos(total / 2), 1j * np.sin(total / 2)], atol=1e-12)


def test_parametric_gate_positions_follow_edits():
    pqc = q.ParametricQuantumCircuit(1)
    g = G.ParametricRX(0, 0.1)

    def positions():
        return [pqc.get_parametric_gate_position(i) for i in range(pqc.get_parameter_count())]

    steps = [
        (lambda: pqc.add_parametric_gate(g), [0]),
        (lambda: pqc.add_parametric_gate(g), [0, 1]),
        (lambda: pqc.add_gate(g), [0, 1]),
        (lambda: pqc.add_parametric_gate(g, 0), [1, 2, 0]),
        (lambda: pqc.add_gate(g, 0), [2, 3, 1]),
        (lambda: pqc.add_parametric_gate(g, 0), [3, 4, 2, 0]),
        (lambda: pqc.remove_gate(4), [3, 2, 0]),
        (lambda: pqc.remove_gate(1), [2, 1, 0]),
        (lambda: pqc.add_parametric_gate(g), [2, 1, 0, 4]),
        (lambda: pqc.add_parametric_gate(g, 2), [3, 1, 0, 5, 2]),
        (lambda: pqc.remove_gate(3), [1, 0, 4, 2]),
    ]
    for act, want in steps:
        act()
        assert positions() == want
    st = q.QuantumState(1)
    pqc.update_quantum_state(st)
    angle = 0.1 * pqc.get_gate_count()
    np.testing.assert_allclose(st.get_vector(), [np.cos(angle / 2), 1j * np.sin(angle / 2)], atol=1e-12)


def test_general_quantum_operator_of_random_unitary():
    n = 4
    a, b, c = q.QuantumState(n), q.QuantumState(n), q.QuantumState(n)
    b.set_Haar_random_state(11)
    u = G.RandomUnitary([0, 1, 2, 3], 12)
    op = q.to_general_quantum_operator(u, n, 1e-5)
    op.apply_to_state(a, b, c)
    u.update_quantum_state(b)
    assert S.inner_product(b, c).real > 0.99
    np.testing.assert_allclose(c.get_vector(), b.get_vector(), atol=1e-5)
