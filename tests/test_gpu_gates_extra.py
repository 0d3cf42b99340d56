"""The less common gate families on the device against numpy: state reflection
(src/csim/update_ops_reflection.cpp:14-27), linear combinations, CP / CPTP maps with
any Kraus gates (branch rules of src/cppsim/gate_general.hpp:380-690, density-matrix
sums over copies beyond three qubits), instruments (classical register writes),
adaptive gates, reversible boolean permutations, FusedSWAP and
multiply_elementwise_function."""
from __future__ import annotations

import numpy as np
import pytest

import qulacs_amd as q
from qulacs_amd import gate

pytestmark = pytest.mark.gpu
TOL = 1e-12
PAULI = [np.eye(2), np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.diag([1.0, -1.0])]


def _sv(n, seed):
    s = q.QuantumState(n)
    s.set_Haar_random_state(seed)
    return s


def _full(m, targets, n):
    """Operator m on `targets` (bit i of m's index <-> targets[i]) as a 2^n matrix."""
    d = 1 << n
    U = np.zeros((d, d), dtype=complex)
    k = len(targets)
    for col in range(d):
        ci = sum(((col >> t) & 1) << i for i, t in enumerate(targets))
        base = col & ~sum(1 << t for t in targets)
        for ri in range(1 << k):
            row = base | sum(((ri >> i) & 1) << t for i, t in enumerate(targets))
            U[row, col] += m[ri, ci]
    return U


def _pauli_full(targets, paulis, n):
    m = np.eye(1, dtype=complex)
    for p in reversed(paulis):
        m = np.kron(m, PAULI[p])
    return _full(m, targets, n)


def test_state_reflection():
    n = 6
    r, s = _sv(n, 1), _sv(n, 2)
    rv, sv = r.get_vector(), s.get_vector()
    g = gate.StateReflection(r)
    g.update_quantum_state(s)
    want = 2 * np.vdot(rv, sv) * rv - sv
    np.testing.assert_allclose(s.get_vector(), want, atol=TOL)
    # reflecting r itself leaves it unchanged; the gate keeps its own copy of r
    r2 = _sv(n, 1)
    r.set_zero_state()
    g.update_quantum_state(r2)
    np.testing.assert_allclose(r2.get_vector(), rv, atol=TOL)
    with pytest.raises(RuntimeError):
        g.update_quantum_state(q.QuantumState(n + 1))
    with pytest.raises(RuntimeError):
        g.update_quantum_state(q.DensityMatrix(n))


def test_linear_combination():
    n = 5
    s = _sv(n, 3)
    v = s.get_vector()
    coefs = [0.3 + 0.1j, -0.7, 2j]
    gates = [gate.X(1), gate.CNOT(0, 3), gate.RY(4, 0.4)]
    g = gate.LinearCombination(coefs, gates)
    g.update_quantum_state(s)
    M = sum(c * _full(np.asarray(x.get_matrix()), x.get_target_index_list(), n) if not x.get_control_index_list()
            else c * _cnot_full(0, 3, n) for c, x in zip(coefs, gates))
    np.testing.assert_allclose(s.get_vector(), M @ v, atol=TOL)
    assert g.get_coef_list() == coefs
    assert [x.get_name() for x in g.get_gate_list()] == ["X", "CNOT", "Y-rotation"]


def _cnot_full(c, t, n):
    d = 1 << n
    U = np.zeros((d, d))
    for x in range(d):
        U[x ^ (1 << t) if (x >> c) & 1 else x, x] = 1
    return U


@pytest.mark.parametrize("sign", [+1, -1])
def test_multi_qubit_pauli_measurement_on_eigenstates(sign):
    n = 4
    targets, paulis = [0, 2, 3], [1, 3, 2]
    P = _pauli_full(targets, paulis, n)
    w, V = np.linalg.eigh(P)
    v = V[:, np.argmax(w) if sign > 0 else np.argmin(w)]
    s = q.QuantumState(n)
    s.load(v)
    g = gate.MultiQubitPauliMeasurement(targets, paulis, 2, 7)
    g.update_quantum_state(s)
    assert s.get_classical_value(2) == (0 if sign > 0 else 1)
    np.testing.assert_allclose(s.get_vector(), v, atol=1e-12)


def test_multi_qubit_pauli_measurement_projects():
    n = 4
    targets, paulis = [1, 3], [3, 1]
    P = _pauli_full(targets, paulis, n)
    for seed in range(6):
        s = _sv(n, 10 + seed)
        v = s.get_vector()
        gate.MultiQubitPauliMeasurement(targets, paulis, 0, seed).update_quantum_state(s)
        out = s.get_classical_value(0)
        proj = (np.eye(1 << n) + (1 - 2 * out) * P) / 2 @ v
        proj /= np.linalg.norm(proj)
        np.testing.assert_allclose(s.get_vector(), proj, atol=1e-12)


def test_cp_branch_rules():
    n = 3
    v0 = _sv(n, 5).get_vector()
    # one branch with total probability 0.01: with probability_normalize it is always taken
    small = gate.DenseMatrix(0, 0.1 * np.eye(2))
    for seed in range(4):
        s = q.QuantumState(n)
        s.load(v0)
        g = gate.CP([small], True, True, True)
        g.set_seed(seed)
        g.update_quantum_state(s)
        np.testing.assert_allclose(s.get_vector(), v0, atol=1e-12)  # rescaled back to norm 1
    # without it the branch is hit only when r < 0.01; otherwise the state is zeroed
    zeros = 0
    for seed in range(20):
        s = q.QuantumState(n)
        s.load(v0)
        g = gate.CP([small], False, False, True)
        g.set_seed(seed)
        g.update_quantum_state(s)
        out = s.get_vector()
        if np.abs(out).max() == 0:
            zeros += 1
        else:
            np.testing.assert_allclose(out, 0.1 * v0, atol=1e-12)  # not renormalised
    assert zeros >= 15
    # projectors without renormalisation: the state is P_i psi for the drawn branch
    s = q.QuantumState(n)
    s.load(v0)
    g = gate.CP([gate.P0(1), gate.P1(1)], False, False, False)
    g.set_seed(3)
    g.update_quantum_state(s)
    out = s.get_vector()
    cands = [np.where(((np.arange(8) >> 1) & 1) == b, v0, 0) for b in (0, 1)]
    assert min(np.abs(out - c).max() for c in cands) < 1e-12


def test_cptp_with_controlled_kraus_on_density_matrix_beyond_three_qubits():
    n = 5
    a = _sv(n, 8).get_vector()
    rho0 = np.outer(a, a.conj())
    d = q.DensityMatrix(n)
    d.load(rho0)
    p = 0.3
    k0 = gate.DenseMatrix([0, 1, 2, 3], np.sqrt(1 - p) * np.eye(16))
    k1 = gate.DenseMatrix(4, np.sqrt(p) * np.array([[0, 1], [1, 0]]))
    k1.add_control_qubit(2, 1)
    k2 = gate.DenseMatrix(4, np.sqrt(p) * np.eye(2))
    k2.add_control_qubit(2, 0)
    ch = gate.CPTP([k0, k1, k2])
    assert ch.get_target_index_list() == [0, 1, 2, 3, 4]
    ch.update_quantum_state(d)
    # a controlled Kraus operator acts as the identity where its control is not met
    K = [np.sqrt(1 - p) * np.eye(32)]
    X4 = _full(PAULI[1], [4], n)
    c1 = np.diag([float((x >> 2) & 1) for x in range(32)])
    c0 = np.eye(32) - c1
    K.append(np.sqrt(p) * X4 @ c1 + c0)
    K.append(np.sqrt(p) * c0 + c1)
    want = sum(k @ rho0 @ k.conj().T for k in K)
    np.testing.assert_allclose(d.get_matrix(), want, atol=1e-12)


def test_probabilistic_instrument_writes_branch():
    n = 3
    for seed in range(5):
        s = q.QuantumState(n)
        s.set_zero_state()
        g = gate.ProbabilisticInstrument([0.3, 0.3, 0.2], [gate.X(0), gate.X(1), gate.X(2)], 1)
        g.set_seed(seed)
        g.update_quantum_state(s)
        k = s.get_classical_value(1)
        v = s.get_vector()
        if k < 3:
            assert abs(v[1 << k]) == pytest.approx(1.0)
        else:
            assert abs(v[0]) == pytest.approx(1.0)  # the implicit identity branch


def test_adaptive():
    s = q.QuantumState(2)
    s.set_zero_state()
    gate.Measurement(0, 0, 1).update_quantum_state(s)  # |00>: register[0] = 0
    flip = gate.Adaptive(gate.X(1), lambda reg: reg[0] == 1)
    flip.update_quantum_state(s)
    assert abs(s.get_vector()[0]) == pytest.approx(1.0)
    s.set_classical_value(0, 1)
    flip.update_quantum_state(s)
    assert abs(s.get_vector()[2]) == pytest.approx(1.0)
    with_id = gate.Adaptive(gate.X(1), lambda reg, i: reg[i] == 1, 0)
    c = with_id.copy()
    c.update_quantum_state(s)
    assert abs(s.get_vector()[0]) == pytest.approx(1.0)


def test_reversible_boolean_and_fused_swap():
    n = 6
    s = _sv(n, 9)
    v = s.get_vector()
    targets = [4, 1, 2]
    g = gate.ReversibleBoolean(targets, lambda x, dim: (x + 3) % dim)
    g.update_quantum_state(s)
    perm = np.zeros((8, 8))
    for x in range(8):
        perm[(x + 3) % 8, x] = 1
    np.testing.assert_allclose(s.get_vector(), _full(perm, targets, n) @ v, atol=TOL)
    with pytest.raises(RuntimeError):
        gate.ReversibleBoolean([0, 1], lambda x, dim: 0)

    s = _sv(n, 10)
    v = s.get_vector()
    fs = gate.FusedSWAP(0, 3, 3)
    assert fs.get_target_index_list() == [0, 1, 2, 3, 4, 5]
    fs.update_quantum_state(s)
    idx = np.arange(1 << n)
    src = ((idx & 7) << 3) | (idx >> 3)
    np.testing.assert_allclose(s.get_vector(), v[src], atol=0)
    np.testing.assert_allclose(_full(np.asarray(fs.get_matrix()), fs.get_target_index_list(), n) @ v, v[src], atol=0)
    d = q.DensityMatrix(n)
    d.load(v)
    gate.FusedSWAP(0, 3, 3).update_quantum_state(d)
    np.testing.assert_allclose(d.get_matrix(), np.outer(v[src], v[src].conj()), atol=TOL)
    with pytest.raises(RuntimeError):
        gate.FusedSWAP(0, 2, 3)


def test_multiply_elementwise_function():
    n = 7
    s = _sv(n, 11)
    v = s.get_vector()
    s.multiply_elementwise_function(lambda i: complex(np.cos(i), 0.5 * i))
    i = np.arange(1 << n)
    np.testing.assert_allclose(s.get_vector(), v * (np.cos(i) + 0.5j * i), atol=1e-12)


def test_noise_simulator():
    n = 3
    c = q.QuantumCircuit(n)
    c.add_H_gate(0)
    c.add_CNOT_gate(0, 1)
    sim = q.NoiseSimulator(c, q.QuantumState(n))
    r = sim.execute_and_get_result(50)
    assert r.get_count() == 1 and r.get_frequency(0) == 50  # no noise: one trajectory
    v = r.get_state(0).get_vector()
    np.testing.assert_allclose(np.abs(v[[0, 3]]) ** 2, [0.5, 0.5], atol=1e-12)
    samples = sim.execute(200)
    assert len(samples) == 200 and set(samples) <= {0, 3}

    c = q.QuantumCircuit(n)
    c.add_gate(gate.BitFlipNoise(0, 0.3))
    c.add_X_gate(2)
    c.add_gate(gate.DepolarizingNoise(1, 0.0))
    sim = q.NoiseSimulator(c)
    sim.set_seed(11)
    r = sim.execute_and_get_result(4000)
    freq = {}
    for i in range(r.get_count()):
        vec = r.get_state(i).get_vector()
        k = int(np.argmax(np.abs(vec)))
        assert abs(vec[k]) == pytest.approx(1.0)
        freq[k] = freq.get(k, 0) + r.get_frequency(i)
    assert set(freq) <= {4, 5} and sum(freq.values()) == 4000
    assert abs(freq.get(5, 0) / 4000 - 0.3) < 0.04
    with pytest.raises(RuntimeError):
        bad = q.QuantumCircuit(1)
        bad.add_gate(gate.AmplitudeDampingNoise(0, 0.1))
        q.NoiseSimulator(bad)


def test_causal_cone_simulator_matches_full_simulation():
    n = 8
    rng = np.random.default_rng(4)
    c = q.ParametricQuantumCircuit(n)
    for layer in range(3):
        for i in range(n):
            c.add_parametric_RY_gate(i, float(rng.uniform(0, 6.28)))
        for i in range(layer % 2, n - 1, 2):
            c.add_CNOT_gate(i, i + 1)
    c.add_gate(gate.PauliRotation([2, 5], [1, 3], 0.3))
    o = q.Observable(n)
    o.add_operator(0.7, "Z 0")
    o.add_operator(-1.2, "X 3 Z 4")
    o.add_operator(0.4, "Y 7")
    o.add_operator(0.25, "")
    s = q.QuantumState(n)
    s.set_zero_state()
    c.update_quantum_state(s)
    want = o.get_expectation_value(s)
    cc = q.CausalConeSimulator(c, o)
    got = cc.get_expectation_value()
    assert abs(got - want) < 1e-10
    circuits = cc.get_circuit_list()
    assert len(circuits) == 4 and circuits[3] == []  # the identity term needs no circuit
    # Z 0 and Y 7 see four qubits each through the three brickwork layers
    assert [x.get_qubit_count() for x in circuits[0]] == [4] and [x.get_qubit_count() for x in circuits[2]] == [4]
    assert cc.get_coef_list() == [0.7, -1.2, 0.4, 0.25]


def test_fused_swap_every_placement_on_10_qubits():
    # test/cppsim/test_gate.cpp:461-473: every (t0, t1, block) whose two blocks fit
    # and do not overlap, against the bit-block permutation of the indices
    n = 10
    idx = np.arange(1 << n)
    st = q.QuantumState(n)
    st.set_Haar_random_state(17)
    v = st.get_vector()
    cases = 0
    for t0 in range(n):
        for t1 in range(n):
            if t0 == t1:
                continue
            for bs in range(1, min(abs(t1 - t0), n - t0, n - t1) + 1):
                mask = (1 << bs) - 1
                lo, hi = (idx >> t0) & mask, (idx >> t1) & mask
                moved = idx & ~((mask << t0) | (mask << t1)) | (lo << t1) | (hi << t0)
                want = np.empty_like(v)
                want[moved] = v
                st.load(v)
                gate.FusedSWAP(t0, t1, bs).update_quantum_state(st)
                np.testing.assert_allclose(st.get_vector(), want, atol=0, err_msg=str((t0, t1, bs)))
                cases += 1
    assert cases == 190
