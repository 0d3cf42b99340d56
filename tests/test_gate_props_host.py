"""Gate class flags and commutation (no device needed): the per-gate values the
reference's constructors set (src/cppsim/gate_named_one.hpp:80-460,
gate_named_two.hpp:80-225, gate_named_pauli.hpp:27-190), the merge rule
(src/cppsim/gate_merge.cpp:19-150, 351), is_commute against the commutator of the
gates' full matrices, to_general_quantum_operator against the gate matrix it
decomposes, and the circuit helpers merge_circuit / add_noise_gate /
add_observable_rotation_gate against matrix exponentials."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

q = pytest.importorskip("qulacs_amd")
from qulacs_amd import GeneralQuantumOperator, Observable, ParametricQuantumCircuit, QuantumCircuit, circuit, gate  # noqa: E402

N = 3


def full(g, n=N):
    """The gate's 2^n x 2^n matrix, controls included (qubit q = index bit q)."""
    m = np.asarray(g.get_matrix())
    t = list(g.get_target_index_list())
    cs = list(zip(g.get_control_index_list(), g.get_control_value_list()))
    d = 1 << n
    U = np.zeros((d, d), dtype=complex)
    for col in range(d):
        if any(((col >> c) & 1) != v for c, v in cs):
            U[col, col] = 1
            continue
        ci = sum(((col >> tq) & 1) << k for k, tq in enumerate(t))
        base = col & ~sum(1 << tq for tq in t)
        for ri in range(1 << len(t)):
            row = base | sum(((ri >> k) & 1) << tq for k, tq in enumerate(t))
            U[row, col] += m[ri, ci]
    return U


# (factory, commutation flags of the target(s), (Pauli, Clifford, Gaussian))
NAMED = [
    (lambda: gate.Identity(0), [7], (1, 1, 1)),
    (lambda: gate.X(0), [1], (1, 1, 0)),
    (lambda: gate.Y(0), [2], (1, 1, 0)),
    (lambda: gate.Z(0), [4], (1, 1, 1)),
    (lambda: gate.H(0), [0], (0, 1, 0)),
    (lambda: gate.S(0), [4], (0, 1, 1)),
    (lambda: gate.Sdag(0), [4], (0, 1, 1)),
    (lambda: gate.T(0), [4], (0, 0, 1)),
    (lambda: gate.Tdag(0), [4], (0, 0, 1)),
    (lambda: gate.sqrtX(0), [1], (0, 1, 0)),
    (lambda: gate.sqrtXdag(0), [1], (0, 1, 0)),
    (lambda: gate.sqrtY(0), [2], (0, 1, 0)),
    (lambda: gate.sqrtYdag(0), [2], (0, 1, 0)),
    (lambda: gate.P0(0), [0], (0, 1, 1)),
    (lambda: gate.P1(0), [0], (0, 1, 1)),
    (lambda: gate.RX(0, 0.3), [1], (0, 0, 0)),
    (lambda: gate.RY(0, 0.3), [2], (0, 0, 0)),
    (lambda: gate.RZ(0, 0.3), [4], (0, 0, 0)),
    (lambda: gate.CNOT(1, 0), [1], (0, 1, 0)),
    (lambda: gate.CZ(1, 0), [4], (0, 1, 0)),
    (lambda: gate.SWAP(0, 1), [0, 0], (0, 1, 0)),
    (lambda: gate.DenseMatrix(0, np.eye(2)), [0], (0, 0, 0)),
    (lambda: gate.DiagonalMatrix([0], [1, 1j]), [0], (0, 0, 0)),
    (lambda: gate.Pauli([0, 1, 2], [1, 2, 3]), [1, 2, 4], (0, 0, 0)),
    (lambda: gate.PauliRotation([0, 2], [3, 0], 0.2), [4, 0], (0, 0, 0)),
    (lambda: gate.ParametricRY(0, 0.1), [2], (0, 0, 0)),
]


def _flags(g):
    # is_diagonal <=> every target commutes with Z; the flags themselves show in to_string
    lines = [l for l in g.to_string().splitlines() if ": commute" in l]
    out = []
    for l in lines:
        rest = l.split(": commute", 1)[1]
        out.append((1 if "X" in rest else 0) | (2 if "Y" in rest else 0) | (4 if "Z" in rest else 0))
    return out


@pytest.mark.parametrize("make,comm,props", NAMED)
def test_named_gate_flags(make, comm, props):
    g = make()
    assert _flags(g) == comm
    assert (g.is_Pauli(), g.is_Clifford(), g.is_Gaussian()) == tuple(bool(x) for x in props)
    assert g.is_diagonal() == all(f & 4 for f in comm)
    assert g.is_parametric() == g.get_name().startswith("Parametric")
    pv = g.get_property_value()
    assert pv & 7 == props[0] | props[1] << 1 | props[2] << 2
    assert bool(pv & 8) == g.is_parametric()


def test_add_control_clears_pauli_and_gaussian():
    g = gate.to_matrix_gate(gate.Z(0))
    assert not g.is_Pauli() and not g.is_Clifford()  # a converted gate has no class flags
    z = gate.Z(0)
    assert _flags(gate.to_matrix_gate(z)) == [4]  # but keeps its commutation
    t = gate.TOFFOLI(0, 1, 2)
    assert _flags(t) == [1] and t.get_control_index_value_list() == [(0, 1), (1, 1)]


def test_merge_rule():
    m = gate.merge(gate.X(0), gate.Z(0))
    assert _flags(m) == [0]
    assert m.is_Pauli() and m.is_Clifford() and not m.is_Gaussian()  # AND of both gates' classes
    # a control shared with the same value stays a control; targets ascend
    m = gate.merge(gate.CNOT(0, 2), gate.CNOT(0, 1))
    assert m.get_control_index_list() == [0] and m.get_target_index_list() == [1, 2]
    assert _flags(m) == [1, 1]
    np.testing.assert_allclose(full(m), full(gate.CNOT(0, 1)) @ full(gate.CNOT(0, 2)), atol=1e-15)
    # control in one gate only: folded into a Z-commuting target
    m = gate.merge(gate.CZ(0, 1), gate.RZ(0, 0.4))
    assert m.get_control_index_list() == [] and _flags(m) == [4, 4] and m.is_diagonal()
    m = gate.merge(gate.H(0), gate.CNOT(0, 1))
    assert _flags(m) == [0, 1]


def test_is_commute_agrees_with_matrices_when_it_claims_commuting():
    makers = [lambda: gate.X(0), lambda: gate.Z(0), lambda: gate.Y(1), lambda: gate.H(2), lambda: gate.S(1),
              lambda: gate.RX(0, 0.7), lambda: gate.RZ(1, 0.2), lambda: gate.CNOT(0, 1), lambda: gate.CZ(1, 2),
              lambda: gate.CNOT(2, 0), lambda: gate.SWAP(0, 2), lambda: gate.Pauli([0, 1], [1, 1]),
              lambda: gate.PauliRotation([1, 2], [3, 3], 0.5), lambda: gate.T(2), lambda: gate.sqrtY(1)]
    for a in makers:
        for b in makers:
            ga, gb = a(), b()
            A, B = full(ga), full(gb)
            if ga.is_commute(gb):
                assert np.abs(A @ B - B @ A).max() < 1e-12, (ga.get_name(), gb.get_name())
            assert ga.is_commute(gb) == gb.is_commute(ga)
    assert gate.CNOT(0, 1).is_commute(gate.X(1))
    assert not gate.CNOT(0, 1).is_commute(gate.Z(1))
    assert gate.CNOT(0, 1).is_commute(gate.Z(0))
    assert not gate.CNOT(0, 1).is_commute(gate.X(0))
    assert gate.CZ(0, 1).is_commute(gate.CZ(1, 2))


def test_optimizer_uses_commutation_flags():
    # X(0), Z(1), X(0): the middle gate commutes with both, so the two X merge
    c = QuantumCircuit(2)
    c.add_X_gate(0)
    c.add_Z_gate(1)
    c.add_X_gate(0)
    opt = circuit.QuantumCircuitOptimizer()
    before = full(gate.merge(gate.merge(c.get_gate(0), c.get_gate(1)), c.get_gate(2)), 2)
    opt.optimize(c, 1)
    assert c.get_gate_count() == 2
    after = np.eye(4, dtype=complex)
    for i in range(c.get_gate_count()):
        after = full(c.get_gate(i), 2) @ after
    np.testing.assert_allclose(after, before, atol=1e-14)


def test_probabilistic_optimize_and_cumulative():
    g = gate.Probabilistic([0.1, 0.6, 0.3], [gate.X(0), gate.Y(0), gate.Z(0)])
    np.testing.assert_allclose(g.get_cumulative_distribution(), [0, 0.1, 0.7, 1.0], atol=1e-15)
    g.optimize_ProbablisticGate()
    assert g.get_distribution() == [0.6, 0.3, 0.1]
    assert [x.get_name() for x in g.get_gate_list()] == ["Y", "Z", "X"]
    np.testing.assert_allclose(g.get_cumulative_distribution(), [0, 0.6, 0.9, 1.0], atol=1e-15)
    assert _flags(g) == [0]  # X & Y & Z
    d = gate.DepolarizingNoise(1, 0.1)
    assert _flags(d) == [0] and not d.is_diagonal()
    assert gate.DephasingNoise(0, 0.1).is_diagonal()


def test_multiply_scalar():
    g = gate.H(0)
    g.multiply_scalar(2j)
    np.testing.assert_allclose(g.get_matrix(), 2j * np.array([[1, 1], [1, -1]]) / np.sqrt(2), atol=1e-15)


def _op_full(o, n):
    return o.get_matrix().toarray()


@pytest.mark.parametrize("k", [1, 2, 3])
def test_to_general_quantum_operator_reconstructs_the_gate(k):
    rng = np.random.default_rng(k)
    targets = list(rng.permutation(4)[:k])
    m = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    g = gate.DenseMatrix(targets, m)
    o = q.to_general_quantum_operator(g, 4, 1e-12)
    np.testing.assert_allclose(_op_full(o, 4), full(g, 4), atol=1e-12)
    assert o.get_term_count() == 4 ** k
    # term order: (row bits, column bits) of the transform; the first term is the identity
    assert o.get_term(0).get_pauli_string() == ""
    o2 = q.to_general_quantum_operator(gate.H(1), 2, 1e-6)
    assert sorted(o2.get_term(i).get_pauli_string() for i in range(o2.get_term_count())) == ["X 1", "Z 1"]
    with pytest.raises(RuntimeError):
        q.to_general_quantum_operator(gate.CNOT(0, 1), 2, 1e-6)


def test_merge_circuit_and_parametric_merge():
    a = ParametricQuantumCircuit(2)
    a.add_H_gate(0)
    a.add_parametric_RX_gate(1, 0.3)
    b = ParametricQuantumCircuit(2)
    b.add_CNOT_gate(0, 1)
    b.add_parametric_RZ_gate(0, 0.7)
    a.merge_circuit(b)
    assert a.get_gate_count() == 4 and a.get_parameter_count() == 2
    assert a.get_parameter(1) == 0.7 and a.get_parametric_gate_position(1) == 3
    c = QuantumCircuit(2)
    c.add_X_gate(0)
    c.merge_circuit(c)
    assert c.get_gate_count() == 2


def test_add_noise_gate():
    c = QuantumCircuit(3)
    c.add_noise_gate(gate.CNOT(0, 1), "Depolarizing", 0.1)
    assert c.get_gate_count() == 2 and c.get_gate(1).get_name() == "Probabilistic"
    assert sorted(c.get_gate(1).get_target_index_list()) == [0, 1]
    c.add_noise_gate(gate.X(2), "AmplitudeDamping", 0.2)
    assert c.get_gate(3).get_name() == "CPTP"
    for t in ("BitFlip", "Dephasing", "IndependentXZ"):
        c.add_noise_gate(gate.H(1), t, 0.05)
    assert c.get_gate_count() == 10
    with pytest.raises(RuntimeError):
        c.add_noise_gate(gate.CNOT(0, 1), "BitFlip", 0.1)
    with pytest.raises(RuntimeError):
        c.add_noise_gate(gate.X(0), "Unknown", 0.1)


def test_observable_rotation_gates_exponentiate():
    n = 3
    o = Observable(n)
    o.add_operator(0.7, "Z 0 Z 1")
    o.add_operator(-0.4, "Z 2")
    H = _op_full(o, n)
    c = QuantumCircuit(n)
    c.add_diagonal_observable_rotation_gate(o, 0.9)
    U = np.eye(1 << n, dtype=complex)
    for i in range(c.get_gate_count()):
        U = full(c.get_gate(i), n) @ U
    # PauliRotation(theta) = exp(+i theta P / 2)
    np.testing.assert_allclose(U, scipy.linalg.expm(0.5j * 0.9 * H), atol=1e-12)
    o.add_operator(0.3, "X 1")
    with pytest.raises(RuntimeError):
        c.add_diagonal_observable_rotation_gate(o, 0.9)
    c2 = QuantumCircuit(n)
    c2.add_observable_rotation_gate(o, 0.5, 200)
    assert c2.get_gate_count() == 3 * 200
    U = np.eye(1 << n, dtype=complex)
    for i in range(c2.get_gate_count()):
        U = full(c2.get_gate(i), n) @ U
    np.testing.assert_allclose(U, scipy.linalg.expm(0.5j * 0.5 * _op_full(o, n)), atol=1e-4)  # Trotter error
    c3 = QuantumCircuit(n)
    c3.add_observable_rotation_gate(o, 0.01, 0)  # repeats = ceil(0.01 * 3 * 100) = 3
    assert c3.get_gate_count() == 3 * 3
    g = GeneralQuantumOperator(n)
    g.add_operator(1j, "X 0")
    with pytest.raises(RuntimeError):
        c3.add_observable_rotation_gate(g, 0.1, 1)


def test_aliases_exist():
    for name in ("ClsOneQubitGate", "ClsOneQubitRotationGate", "ClsOneControlOneTargetGate", "ClsTwoQubitGate",
                 "ClsPauliRotationGate", "QuantumGateSparseMatrix", "QuantumGate_Instrument"):
        assert hasattr(q, name)
    assert q.check_build_for_mpi() is False
    o = Observable(1)
    o.add_operator_move(q.PauliOperator("Z 0", 1.0))
    assert o.get_term_count() == 1


@pytest.mark.parametrize("first,second,controls,targets", [
    (lambda: gate.X(0), lambda: gate.CNOT(0, 1), [], [0, 1]),
    (lambda: gate.CNOT(0, 1), lambda: gate.X(0), [], [0, 1]),
    (lambda: gate.X(1), lambda: gate.CNOT(0, 1), [], [0, 1]),
    (lambda: gate.CNOT(0, 1), lambda: gate.X(1), [], [0, 1]),
    (lambda: gate.CNOT(0, 1), lambda: gate.CZ(0, 1), [0], [1]),
    (lambda: gate.CNOT(0, 1), lambda: gate.CZ(1, 0), [], [0, 1]),
    (lambda: gate.X(2), lambda: gate.CNOT(0, 1), [], [0, 1, 2]),
])
def test_control_merge_cases(first, second, controls, targets):
    """test/cppsim/test_gate.cpp:979-1140: the merged gate's own matrix, in
    ascending target order, is the product of the two full matrices; CNOT then CZ
    on the same pair keeps the shared control and leaves iY on the target."""
    a, b = first(), second()
    m = gate.merge(a, b)
    assert m.get_control_index_list() == controls and m.get_target_index_list() == targets
    n = len(targets) + len(controls)
    want = full(b, n) @ full(a, n)
    if controls:
        np.testing.assert_allclose(m.get_matrix(), 1j * np.array([[0, -1j], [1j, 0]]), atol=1e-14)
    else:
        np.testing.assert_allclose(m.get_matrix(), want, atol=1e-14)
    np.testing.assert_allclose(full(m, n), want, atol=1e-14)


def test_duplicate_qubits_are_rejected():
    """test/cppsim/test_gate.cpp:1540-1625: a gate given the same qubit twice is an
    error; the same factories with distinct qubits build."""
    from scipy.sparse import identity

    good = [lambda: gate.CNOT(10, 13), lambda: gate.CZ(10, 13), lambda: gate.SWAP(10, 13),
            lambda: gate.Pauli([2, 1, 0, 3, 7, 9, 4], [0] * 7),
            lambda: gate.PauliRotation([2, 1, 0, 3, 7, 9, 4], [0] * 7, 0.2),
            lambda: gate.DenseMatrix([10, 13], np.eye(4)), lambda: gate.SparseMatrix([10, 13], identity(4)),
            lambda: gate.DiagonalMatrix([10, 13], np.ones(4)), lambda: gate.RandomUnitary([10, 13]),
            lambda: gate.ParametricPauliRotation([1, 2], [1, 3], 0.1), lambda: gate.TOFFOLI(1, 2, 3),
            lambda: gate.FREDKIN(1, 2, 3)]
    bad = [lambda: gate.CNOT(21, 21), lambda: gate.CZ(21, 21), lambda: gate.SWAP(21, 21),
           lambda: gate.Pauli([0, 1, 3, 1, 5, 6, 2], [0] * 7),
           lambda: gate.PauliRotation([0, 1, 3, 1, 5, 6, 2], [0] * 7, 0.2),
           lambda: gate.DenseMatrix([21, 21], np.eye(4)), lambda: gate.SparseMatrix([21, 21], identity(4)),
           lambda: gate.DiagonalMatrix([21, 21], np.ones(4)), lambda: gate.RandomUnitary([3, 3]),
           lambda: gate.ParametricPauliRotation([2, 2], [1, 3], 0.1), lambda: gate.TOFFOLI(1, 1, 3),
           lambda: gate.FREDKIN(1, 3, 3)]
    for make in good:
        assert make() is not None
    for i, make in enumerate(bad):
        with pytest.raises(Exception, match="must differ|duplicated"):
            make()
