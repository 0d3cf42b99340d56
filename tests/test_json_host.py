"""JSON wire format of gates, circuits and operators (no device needed): the
property-tree layout of the reference's to_ptree (every leaf a string, arrays of
unnamed children) and round trips through from_json, as
python/tests/single_cpu/test_json.py does for the reference; states are covered
by tests/test_gpu_json.py."""
from __future__ import annotations

import json
import pickle

import numpy as np
import pytest

q = pytest.importorskip("qulacs_amd")
from qulacs_amd import (  # noqa: E402
    GeneralQuantumOperator,
    Observable,
    ParametricQuantumCircuit,
    PauliOperator,
    QuantumCircuit,
    circuit,
    gate,
    observable,
    quantum_operator,
)


def _gates(rng):
    m = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    return [
        gate.Identity(0), gate.X(0), gate.Y(1), gate.Z(2), gate.H(0), gate.S(1), gate.Sdag(2), gate.T(0),
        gate.Tdag(1), gate.sqrtX(2), gate.sqrtXdag(0), gate.sqrtY(1), gate.sqrtYdag(2), gate.P0(0), gate.P1(1),
        gate.CNOT(0, 1), gate.CZ(2, 0), gate.SWAP(1, 2),
        gate.Pauli([0, 1], [1, 2]), gate.PauliRotation([0, 2], [3, 1], rng.uniform()),
        gate.RX(0, rng.uniform()), gate.RY(1, rng.uniform()), gate.RZ(2, rng.uniform()),
        gate.RotX(0, rng.uniform()), gate.U1(0, rng.uniform()), gate.U2(1, rng.uniform(), rng.uniform()),
        gate.U3(2, rng.uniform(), rng.uniform(), rng.uniform()),
        gate.merge(gate.X(0), gate.Y(1)), gate.to_matrix_gate(gate.X(0)), gate.DenseMatrix(1, m),
        gate.DenseMatrix([0, 2], np.kron(m, m)), gate.DiagonalMatrix([0, 1], np.exp(1j * rng.uniform(size=4))),
        gate.RandomUnitary([0, 1, 2]),
    ]


def test_gate_round_trip_keeps_the_matrix():
    rng = np.random.default_rng(11)
    for g in _gates(rng):
        s = g.to_json()
        d = json.loads(s)
        assert isinstance(d["name"], str) and d["name"].endswith("Gate"), s
        g2 = gate.from_json(s)
        assert g2.get_target_index_list() == g.get_target_index_list(), d["name"]
        assert g2.get_control_index_list() == g.get_control_index_list(), d["name"]
        assert np.array_equal(g2.get_matrix(), g.get_matrix()), d["name"]


def test_controlled_dense_gate_round_trip():
    g = gate.X(2)
    g.add_control_qubit(0, 1)
    g.add_control_qubit(1, 0)
    d = json.loads(g.to_json())
    assert d["name"] == "DenseMatrixGate"
    assert d["control_qubit_list"] == [{"index": "0", "value": "1"}, {"index": "1", "value": "0"}]
    g2 = gate.from_json(g.to_json())
    assert g2.get_control_index_list() == [0, 1] and g2.get_control_value_list() == [1, 0]
    assert np.array_equal(g2.get_matrix(), g.get_matrix())


def test_reference_layout_of_named_and_matrix_gates():
    """Keys, nesting and string-valued leaves as the reference's to_ptree
    (gate_named_one.hpp:331-336, gate_named_two.hpp:231-237, gate_matrix.cpp:312-319,
    utility.cpp:126-185)."""
    assert json.loads(gate.X(3).to_json()) == {"name": "XGate", "target_qubit": "3"}
    assert json.loads(gate.P0(1).to_json()) == {"name": "Projection-0Gate", "target_qubit": "1"}
    assert json.loads(gate.CNOT(0, 2).to_json()) == {"name": "CNOTGate", "control_qubit": "0", "target_qubit": "2"}
    assert json.loads(gate.SWAP(1, 2).to_json()) == {"name": "SWAPGate", "target_qubit_list": ["1", "2"]}
    assert json.loads(gate.RX(0, 0.5).to_json()) == {"name": "X-rotationGate", "target_qubit": "0", "angle": "0.5"}
    d = json.loads(gate.DenseMatrix(0, np.array([[0, 1j], [2, 3]])).to_json())
    assert d["name"] == "DenseMatrixGate" and d["target_qubit_list"] == ["0"] and d["control_qubit_list"] == ""
    assert d["matrix"]["rows"] == "2" and d["matrix"]["cols"] == "2"
    assert d["matrix"]["data"][1] == {"real": "0", "imag": "1"}  # row-major
    p = json.loads(gate.Pauli([0, 1], [1, 3]).to_json())
    assert p["name"] == "PauliGate" and p["pauli"]["name"] == "PauliOperator"
    assert p["pauli"]["pauli_list"][1] == {"name": "SinglePauliOperator", "index": "1", "pauli_id": "3"}


def test_reader_accepts_reference_documents():
    """Plain JSON numbers and booleans (as another writer may emit them) and the
    reference's sparse-matrix gate (loaded as a dense matrix)."""
    g = gate.from_json('{"name": "Y-rotationGate", "target_qubit": 2, "angle": 0.25}')
    assert np.array_equal(g.get_matrix(), gate.RY(2, 0.25).get_matrix())
    sparse = {"name": "SparseMatrixGate", "target_qubit_list": ["0"], "control_qubit_list": "",
              "matrix": {"rows": "2", "cols": "2", "data": [
                  {"row": "0", "col": "1", "value": {"real": "1", "imag": "0"}},
                  {"row": "1", "col": "0", "value": {"real": "1", "imag": "0"}}]}}
    g = gate.from_json(json.dumps(sparse))
    assert np.array_equal(g.get_matrix(), gate.X(0).get_matrix())
    prob = {"name": "ProbabilisticGate", "distribution": ["0.25", "0.75"],
            "gate_list": [{"name": "XGate", "target_qubit": "0"}, {"name": "IGate", "target_qubit": "0"}],
            "is_instrument": False}
    g = gate.from_json(json.dumps(prob))
    assert g.get_distribution() == [0.25, 0.75]
    with pytest.raises(RuntimeError):
        gate.from_json('{"name": "NoSuchGate", "target_qubit": "0"}')
    with pytest.raises(RuntimeError):
        gate.from_json('{"name": "XGate"}')
    with pytest.raises(RuntimeError):
        gate.from_json('{"name": "XGate", "target_qubit": "0"')


def test_parametric_gates_round_trip():
    for g in [gate.ParametricRX(0, 0.1), gate.ParametricRY(1, 0.2), gate.ParametricRZ(2, 0.3),
              gate.ParametricPauliRotation([0, 1], [1, 1], 0.4)]:
        g2 = gate.from_json(g.to_json())
        assert g2.is_parametric() and g2.get_parameter_value() == g.get_parameter_value()
        assert np.array_equal(g2.get_matrix(), g.get_matrix())


def test_noise_and_channels_round_trip():
    r = 0.3
    for g in [gate.Probabilistic([r, 1 - r], [gate.X(0), gate.Y(0)]), gate.BitFlipNoise(0, 0.1),
              gate.DephasingNoise(0, 0.2), gate.IndependentXZNoise(0, 0.05), gate.DepolarizingNoise(0, 0.07),
              gate.TwoQubitDepolarizingNoise(0, 1, 0.09)]:
        g2 = gate.from_json(g.to_json())
        assert np.allclose(g2.get_distribution(), g.get_distribution(), rtol=0, atol=0)
        assert np.array_equal(g2.get_superoperator(), g.get_superoperator())
    for g in [gate.AmplitudeDampingNoise(0, 0.3), gate.CPTP([gate.P0(0), gate.P1(0)])]:
        g2 = gate.from_json(g.to_json())
        a, b = g.get_gate_list(), g2.get_gate_list()
        assert [x.get_name() for x in a] == [x.get_name() for x in b]
        for x, y in zip(a, b):
            assert np.array_equal(x.get_matrix(), y.get_matrix())


def test_circuits_round_trip_and_pickle():
    c = QuantumCircuit(3)
    for _ in range(3):
        c.add_gate(gate.RandomUnitary([0, 1, 2]))
    c.add_CNOT_gate(0, 2)
    c.add_gate(gate.DepolarizingNoise(1, 0.01))
    for c2 in (circuit.from_json(c.to_json()), pickle.loads(pickle.dumps(c))):
        assert c2.get_gate_count() == c.get_gate_count() and c2.get_qubit_count() == 3
        for i in range(3):
            assert np.array_equal(c2.get_gate(i).get_matrix(), c.get_gate(i).get_matrix())
    p = ParametricQuantumCircuit(2)
    p.add_parametric_RX_gate(0, 0.3)
    p.add_H_gate(1)
    p.add_parametric_multi_Pauli_rotation_gate([0, 1], [1, 1], 0.7)
    p.add_CNOT_gate(0, 1)
    p.add_parametric_RY_gate(1, 0.2)
    assert json.loads(p.to_json())["name"] == "ParametricQuantumCircuit"
    for p2 in (circuit.from_json(p.to_json()), pickle.loads(pickle.dumps(p))):
        assert isinstance(p2, ParametricQuantumCircuit)
        assert p2.get_parameter_count() == 3 and p2.get_gate_count() == 5
        assert [p2.get_parameter(i) for i in range(3)] == [0.3, 0.7, 0.2]
        assert [p2.get_parametric_gate_position(i) for i in range(3)] == [0, 2, 4]


def test_operators_round_trip():
    rng = np.random.default_rng(5)
    op = GeneralQuantumOperator(4)
    for _ in range(5):
        p = PauliOperator(complex(rng.normal(), rng.normal()))
        for _ in range(3):
            p.add_single_Pauli(int(rng.integers(4)), int(rng.integers(4)))
        op.add_operator(p)
    d = json.loads(op.to_json())
    assert d["name"] == "GeneralQuantumOperator" and d["qubit_count"] == "4" and len(d["operator_list"]) == 5
    op2 = quantum_operator.from_json(op.to_json())
    assert np.array_equal(op2.get_matrix().toarray(), op.get_matrix().toarray())
    obs = Observable(3)
    obs.add_operator(0.5, "X 0 Z 2")
    obs.add_operator(-1.25, "Y 1")
    obs2 = observable.from_json(obs.to_json())
    assert isinstance(obs2, Observable)
    assert np.array_equal(obs2.get_matrix().toarray(), obs.get_matrix().toarray())
    bad = GeneralQuantumOperator(1)
    bad.add_operator(1j, "X 0")
    with pytest.raises(RuntimeError):
        observable.from_json(bad.to_json())


def test_toffoli_fredkin_add_sparse():
    """TOFFOLI / FREDKIN as controlled matrix gates (python/cppsim_wrapper.cpp:910-933),
    add as the sum of the gates' matrices over the union of their qubits with shared
    controls kept (src/cppsim/gate_merge.cpp:372-430), SparseMatrix through its dense form."""
    import scipy.sparse as sp

    t = gate.TOFFOLI(0, 1, 2)
    assert t.get_target_index_list() == [2] and t.get_control_index_list() == [0, 1]
    assert t.get_control_value_list() == [1, 1] and np.array_equal(t.get_matrix(), [[0, 1], [1, 0]])
    f = gate.FREDKIN(0, 1, 2)
    assert f.get_target_index_list() == [1, 2] and f.get_control_index_list() == [0]
    assert np.array_equal(f.get_matrix(), gate.SWAP(1, 2).get_matrix())
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0 + 0j, -1.0])
    I2 = np.eye(2)
    assert np.array_equal(gate.add(gate.X(0), gate.Y(0)).get_matrix(), X + Y)
    s = gate.add(gate.X(0), gate.Z(1))
    assert s.get_target_index_list() == [0, 1]
    assert np.array_equal(s.get_matrix(), np.kron(I2, X) + np.kron(Z, I2))  # bit 0 <-> qubit 0
    a, b = gate.X(2), gate.Z(2)
    a.add_control_qubit(0, 1)
    b.add_control_qubit(0, 1)
    c = gate.add(a, b)
    assert c.get_control_index_list() == [0] and np.array_equal(c.get_matrix(), X + Z)
    assert np.array_equal(gate.add([gate.X(0), gate.Y(0), gate.Z(0)]).get_matrix(), X + Y + Z)
    m = np.array([[0.5, 0.25j], [0, 1.0]])
    g = gate.SparseMatrix([1], sp.csc_matrix(m))
    assert g.get_target_index_list() == [1] and np.array_equal(g.get_matrix(), m)
    for g in (t, f, c):
        g2 = gate.from_json(g.to_json())
        assert g2.get_control_index_list() == g.get_control_index_list()
        assert np.array_equal(g2.get_matrix(), g.get_matrix())


def test_instrument_round_trip():
    m = gate.Measurement(1, 3)
    d = json.loads(m.to_json())
    assert d["name"] == "CPTPMapGate" and d["is_instrument"] == "true" and d["classical_register_address"] == "3"
    m2 = gate.from_json(m.to_json())
    assert m2.is_instrument() and len(m2.get_gate_list()) == 2
    assert not gate.AmplitudeDampingNoise(0, 0.1).is_instrument()


def test_newer_gate_kinds_round_trip():
    # CPMapGate (gate_general.hpp:676-690), ProbabilisticGate with a register,
    # LinearCombinationGate, FusedSWAPGate, NoisyEvolution(Fast)Gate
    cp = gate.CP([gate.P0(1), gate.P1(1)], False, True, True)
    d = json.loads(cp.to_json())
    assert d["name"] == "CPMapGate" and d["state_normalize"] == "false" and d["probability_normalize"] == "true"
    cp2 = gate.from_json(cp.to_json())
    assert json.loads(cp2.to_json()) == d
    pi = gate.ProbabilisticInstrument([0.25, 0.75], [gate.X(0), gate.Z(0)], 4)
    d = json.loads(pi.to_json())
    assert d["is_instrument"] == "true" and d["classical_register_address"] == "4"
    pi2 = gate.from_json(pi.to_json())
    assert pi2.is_instrument() and pi2.get_distribution() == [0.25, 0.75]
    lc = gate.LinearCombination([0.5 + 0.25j, -1.0], [gate.X(0), gate.CNOT(0, 2)])
    d = json.loads(lc.to_json())
    assert d["name"] == "LinearCombinationGate" and d["coefs"] == ["(0.5,0.25)", "(-1,0)"]
    lc2 = gate.from_json(lc.to_json())
    assert lc2.get_coef_list() == [0.5 + 0.25j, -1.0]
    assert np.array_equal(lc2.get_matrix(), lc.get_matrix())
    fs = gate.FusedSWAP(1, 4, 2)
    fs2 = gate.from_json(fs.to_json())
    assert fs2.get_target_index_list() == [1, 2, 4, 5] and np.array_equal(fs2.get_matrix(), fs.get_matrix())
    h = Observable(2)
    h.add_operator(0.5, "Z 0 Z 1")
    c = GeneralQuantumOperator(2)
    c.add_operator(0.5, "X 1")
    c.add_operator(0.5j, "Y 1")
    for g in (gate.NoisyEvolution(h, [c], 0.7, 0.01), gate.NoisyEvolution_fast(h, [c], 0.7)):
        g2 = gate.from_json(g.to_json())
        assert g2.get_name() == g.get_name()
        assert json.loads(g2.to_json()) == json.loads(g.to_json())
