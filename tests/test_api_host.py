"""Host-side qulacs API (qulacs_amd._core) without a GPU: gate matrices and
conventions, channel superoperators, circuit and parameter bookkeeping, Pauli
strings and operator matrices.  Nothing here allocates a device state.

Conventions checked against the reference:
  * RX/RY/RZ(theta) = exp(+i theta P / 2) (src/csim/update_ops_pauli_single.cpp:54-116),
    RotX/RotY/RotZ use -theta (src/cppsim/gate_factory.cpp:152-165)
  * U1/U2/U3 are the IBMQ matrices (src/cppsim/gate_factory.cpp:169-181)
  * DepolarizingNoise = Probabilistic({p/3, p/3, p/3, 1-p}, {X, Z, Y, I})
    (src/cppsim/gate_factory.cpp:454-467)
  * operator matrices: python/tests/single_cpu/test_observable.py
"""
from __future__ import annotations

import numpy as np
import pytest

import qulacs_amd as q
from qulacs_amd import gate

I2 = np.eye(2)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]])
Z = np.diag([1.0, -1.0]).astype(complex)
P = {0: I2, 1: X, 2: Y, 3: Z}


def expm_pauli(theta, p):
    """exp(i theta p / 2) for a Pauli product p (p @ p = I)."""
    return np.cos(theta / 2) * np.eye(p.shape[0]) + 1j * np.sin(theta / 2) * p


def kron_targets(mats):
    """Matrix over targets [t0, t1, ...] with bit i <-> mats[i]."""
    out = np.array([[1.0 + 0j]])
    for m in mats:
        out = np.kron(m, out)
    return out


@pytest.mark.parametrize("theta", [0.0, 0.3, -1.7, np.pi])
def test_rotation_conventions(theta):
    for f, p in ((gate.RX, X), (gate.RY, Y), (gate.RZ, Z)):
        assert np.allclose(f(0, theta).get_matrix(), expm_pauli(theta, p), atol=1e-15)
    for f, p in ((gate.RotX, X), (gate.RotY, Y), (gate.RotZ, Z)):
        assert np.allclose(f(0, theta).get_matrix(), expm_pauli(-theta, p), atol=1e-15)
    for f, p in ((gate.RotInvX, X), (gate.RotInvY, Y), (gate.RotInvZ, Z)):
        assert np.allclose(f(0, theta).get_matrix(), expm_pauli(theta, p), atol=1e-15)
    assert np.allclose(gate.ParametricRX(0, theta).get_matrix(), expm_pauli(theta, X))
    assert np.allclose(
        gate.PauliRotation([0, 2], [1, 3], theta).get_matrix(), expm_pauli(theta, kron_targets([X, Z]))
    )


def test_named_gate_identities():
    m = lambda g: g.get_matrix()  # noqa: E731
    assert np.allclose(m(gate.H(0)) @ m(gate.H(0)), I2)
    assert np.allclose(m(gate.S(0)) @ m(gate.S(0)), Z)
    assert np.allclose(m(gate.T(0)) @ m(gate.T(0)), m(gate.S(0)))
    assert np.allclose(m(gate.sqrtX(0)) @ m(gate.sqrtX(0)), X)
    assert np.allclose(m(gate.sqrtY(0)) @ m(gate.sqrtY(0)), Y)
    for a, b in ((gate.S, gate.Sdag), (gate.T, gate.Tdag), (gate.sqrtX, gate.sqrtXdag), (gate.sqrtY, gate.sqrtYdag)):
        assert np.allclose(m(a(0)) @ m(b(0)), I2)
    assert np.allclose(m(gate.P0(0)) + m(gate.P1(0)), I2)
    assert np.allclose(m(gate.SWAP(0, 1)), np.eye(4)[[0, 2, 1, 3]])
    g = gate.CNOT(2, 5)
    assert g.get_target_index_list() == [5] and g.get_control_index_list() == [2]
    assert np.allclose(m(g), X)
    assert np.allclose(m(gate.CZ(0, 1)), Z)


def test_u_gates_are_ibmq_matrices():
    t, ph, la = 0.7, -0.4, 1.9
    want = np.array(
        [[np.cos(t / 2), -np.exp(1j * la) * np.sin(t / 2)],
         [np.exp(1j * ph) * np.sin(t / 2), np.exp(1j * (ph + la)) * np.cos(t / 2)]]
    )
    assert np.allclose(gate.U3(0, t, ph, la).get_matrix(), want)
    assert np.allclose(gate.U1(0, la).get_matrix(), np.diag([1, np.exp(1j * la)]))
    assert np.allclose(gate.U2(0, ph, la).get_matrix(), gate.U3(0, np.pi / 2, ph, la).get_matrix())


def test_inverse_and_merge():
    rng = np.random.default_rng(5)
    for g in (gate.RX(1, 0.4), gate.T(0), gate.RandomUnitary([0, 2], 7), gate.PauliRotation([0, 1], [2, 1], 0.9)):
        m = g.get_matrix()
        assert np.allclose(g.get_inverse().get_matrix() @ m, np.eye(m.shape[0]))
    u = gate.RandomUnitary([0, 1, 3], int(rng.integers(1 << 30))).get_matrix()
    assert np.allclose(u.conj().T @ u, np.eye(8), atol=1e-12)
    # merge(first, second) = second * first over the ascending union of targets
    a, b = gate.H(2), gate.RY(0, 0.3)
    mg = gate.merge(a, b)
    assert mg.get_target_index_list() == [0, 2]
    assert np.allclose(mg.get_matrix(), kron_targets([b.get_matrix(), I2]) @ kron_targets([I2, a.get_matrix()]))
    mg2 = gate.merge([gate.X(0), gate.Y(0), gate.Z(0)])
    assert np.allclose(mg2.get_matrix(), Z @ Y @ X)


def test_dense_and_diagonal_factories():
    rng = np.random.default_rng(2)
    mat = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    g = gate.DenseMatrix([3, 1], mat)
    assert g.get_target_index_list() == [3, 1] and np.allclose(g.get_matrix(), mat)
    assert gate.DenseMatrix(2, mat[:2, :2]).get_target_index_list() == [2]
    d = np.exp(1j * rng.normal(size=4))
    assert np.allclose(gate.DiagonalMatrix([0, 1], d).get_matrix(), np.diag(d))
    with pytest.raises(RuntimeError):
        gate.DenseMatrix([0, 0], mat)
    with pytest.raises(RuntimeError):
        gate.DenseMatrix([0], mat)
    g = gate.to_matrix_gate(gate.CNOT(0, 1))
    g.add_control_qubit(4, 0)
    assert g.get_control_index_list() == [0, 4] and g.get_control_value_list() == [1, 0]


def superop_of(kraus_with_weights):
    """sum_i w_i K_i (x) conj(K_i) with member index = row bits * 2^k + column bits."""
    d = kraus_with_weights[0][1].shape[0]
    s = np.zeros((d * d, d * d), dtype=complex)
    for w, k in kraus_with_weights:
        s += w * np.kron(k, k.conj())
    return s


@pytest.mark.parametrize("p", [0.0, 0.05, 0.3])
def test_noise_superoperators(p):
    dep = gate.DepolarizingNoise(0, p).get_superoperator()
    a, b, c = 1 - 2 * p / 3, 2 * p / 3, 1 - 4 * p / 3
    closed = np.array([[a, 0, 0, b], [0, c, 0, 0], [0, 0, c, 0], [b, 0, 0, a]])
    assert np.allclose(dep, closed, atol=1e-15)
    assert np.allclose(gate.BitFlipNoise(1, p).get_superoperator(), superop_of([(1 - p, I2), (p, X)]))
    assert np.allclose(gate.DephasingNoise(1, p).get_superoperator(), superop_of([(1 - p, I2), (p, Z)]))
    p1, p2 = p * (1 - p), p * p
    assert np.allclose(
        gate.IndependentXZNoise(0, p).get_superoperator(),
        superop_of([(1 - 2 * p1 - p2, I2), (p1, X), (p1, Z), (p2, Y)]),
    )
    k0 = np.array([[1, 0], [0, np.sqrt(1 - p)]])
    k1 = np.array([[0, np.sqrt(p)], [0, 0]])
    assert np.allclose(gate.AmplitudeDampingNoise(0, p).get_superoperator(), superop_of([(1, k0), (1, k1)]))
    two = gate.TwoQubitDepolarizingNoise(3, 1, p)
    assert two.get_target_index_list() == [1, 3]
    terms = [(1 - p if i == 0 else p / 15, kron_targets([P[i % 4], P[i // 4]])) for i in range(16)]
    # the gate acts on (3, 1); its superoperator is over the ascending list (1, 3)
    swap = np.eye(4)[[0, 2, 1, 3]]
    assert np.allclose(two.get_superoperator(), superop_of([(w, swap @ k @ swap) for w, k in terms]))


def test_probabilistic_and_cptp_targets():
    g = gate.Probabilistic([0.2, 0.3], [gate.X(4), gate.CNOT(1, 2)])
    assert g.get_target_index_list() == [1, 2, 4]
    assert np.isclose(sum(g.get_distribution()), 0.5)
    c = gate.CPTP([gate.DenseMatrix(0, np.diag([1, 0.5])), gate.DenseMatrix(0, np.array([[0, np.sqrt(0.75)], [0, 0]]))])
    s = c.get_superoperator()
    # trace preservation: sum_i K_i^dagger K_i = I  <=>  rows y==x of S sum to delta
    d = 2
    tp = np.zeros((d, d), dtype=complex)
    for y in range(d):
        for x in range(d):
            for oy in range(d):
                tp[y, x] += s[oy * d + oy, y * d + x]
    assert np.allclose(tp, np.eye(2))
    with pytest.raises(RuntimeError):
        gate.Probabilistic([0.7, 0.6], [gate.X(0), gate.Z(0)])


def test_circuit_bookkeeping():
    c = q.QuantumCircuit(4)
    c.add_H_gate(0)
    c.add_CNOT_gate(0, 1)
    c.add_RX_gate(2, 0.1)
    c.add_gate(gate.CZ(1, 3))
    c.add_gate(gate.X(3), 0)
    assert c.get_gate_count() == 5 and c.calculate_depth() == 3
    assert c.get_gate(0).get_name() == "X"
    c.remove_gate(0)
    assert c.get_gate(0).get_name() == "H"
    inv = c.get_inverse()
    assert [inv.get_gate(i).get_name() for i in range(inv.get_gate_count())][0] == "CZ"
    with pytest.raises(RuntimeError):
        c.add_X_gate(4)
    with pytest.raises(RuntimeError):
        c.add_gate(gate.CNOT(5, 0))
    # the positional overload checks the same indices and leaves the circuit as it was
    before = [c.get_gate(i).get_name() for i in range(c.get_gate_count())]
    with pytest.raises(RuntimeError):
        c.add_gate(gate.X(4), 0)
    with pytest.raises(RuntimeError):
        c.add_gate(gate.CNOT(7, 1), 1)
    assert [c.get_gate(i).get_name() for i in range(c.get_gate_count())] == before
    assert c.calculate_depth() == 3
    cc = c.copy()
    cc.add_Z_gate(0)
    assert cc.get_gate_count() == c.get_gate_count() + 1


def test_parametric_circuit_bookkeeping():
    c = q.ParametricQuantumCircuit(3)
    c.add_parametric_RX_gate(0, 0.1)
    c.add_H_gate(1)
    c.add_parametric_RY_gate(1, 0.2)
    c.add_parametric_multi_Pauli_rotation_gate([0, 2], [1, 3], 0.3)
    c.add_gate(gate.ParametricRZ(2, 0.9))  # a plain gate, not a parameter
    assert c.get_parameter_count() == 3
    assert [c.get_parametric_gate_position(i) for i in range(3)] == [0, 2, 3]
    c.add_gate(gate.X(0), 1)
    assert [c.get_parametric_gate_position(i) for i in range(3)] == [0, 3, 4]
    with pytest.raises(RuntimeError):
        c.add_gate(gate.X(3), 1)  # qubit 3 of a 3-qubit circuit: rejected, positions unchanged
    with pytest.raises(RuntimeError):
        c.add_parametric_gate(gate.ParametricRX(5, 0.1), 0)
    assert [c.get_parametric_gate_position(i) for i in range(3)] == [0, 3, 4] and c.get_gate_count() == 6
    c.set_parameter(1, 1.5)
    assert c.get_parameter(1) == 1.5 and c.get_gate(3).get_parameter_value() == 1.5
    c.remove_gate(0)
    assert c.get_parameter_count() == 2
    assert [c.get_parametric_gate_position(i) for i in range(2)] == [2, 3]
    assert c.copy().get_parameter_count() == 2
    with pytest.raises(RuntimeError):
        c.get_parameter(2)


def test_pauli_strings_and_operator_matrix():
    p = q.PauliOperator("X 0 y 3 Z 1 I 2", 0.5)
    assert p.get_index_list() == [0, 3, 1] and p.get_pauli_id_list() == [1, 2, 3]
    assert p.get_pauli_string() == "X 0 Y 3 Z 1"
    with pytest.raises(RuntimeError):
        q.PauliOperator("W 0")
    # python/tests/single_cpu/test_observable.py
    obs = q.Observable(3)
    obs.add_operator(0.5, "Z 2")
    obs.add_operator(1.0, "X 0 X 1 X 2")
    obs.add_operator(1.0, "Y 1")
    ans = np.array(
        [
            [0.5, 0, -1j, 0, 0, 0, 0, 1],
            [0, 0.5, 0, -1j, 0, 0, 1, 0],
            [1j, 0, 0.5, 0, 0, 1, 0, 0],
            [0, 1j, 0, 0.5, 1, 0, 0, 0],
            [0, 0, 0, 1, -0.5, 0, -1j, 0],
            [0, 0, 1, 0, 0, -0.5, 0, -1j],
            [0, 1, 0, 0, 1j, 0, -0.5, 0],
            [1, 0, 0, 0, 0, 1j, 0, -0.5],
        ],
        dtype=np.complex128,
    )
    assert np.linalg.norm(ans - obs.get_matrix().todense()) <= 1e-6
    gen = q.GeneralQuantumOperator(3)
    gen.add_operator(0.5j, "Z 2")
    gen.add_operator(1.0, "X 0 X 1 X 2")
    gen.add_operator(1.0, "Y 1")
    ans2 = ans.copy()
    ans2[np.arange(4), np.arange(4)] = 0.5j
    ans2[np.arange(4, 8), np.arange(4, 8)] = -0.5j
    assert np.linalg.norm(ans2 - gen.get_matrix().todense()) <= 1e-6
    assert obs.is_hermitian() and not gen.is_hermitian()
    with pytest.raises(RuntimeError):
        obs.add_operator(1j, "X 0")
    with pytest.raises(RuntimeError):
        obs.add_operator(1.0, "X 3")


def test_aliases():
    assert q.QuantumState is q.QuantumStateMI355 and q.QuantumStateGpu is q.QuantumStateMI355
    assert q.DensityMatrix is q.DensityMatrixMI355
    from qulacs_amd.state import inner_product  # noqa: F401
    import qulacs_amd.gate as g2

    assert g2 is gate


def _random_circuit(rng, n, count):
    c = q.QuantumCircuit(n)
    for _ in range(count):
        k = int(rng.integers(6))
        t = [int(x) for x in rng.choice(n, 2, replace=False)]
        if k == 0:
            c.add_gate(gate.RandomUnitary([t[0]], int(rng.integers(1 << 30))))
        elif k == 1:
            c.add_CNOT_gate(t[0], t[1])
        elif k == 2:
            c.add_CZ_gate(t[0], t[1])
        elif k == 3:
            c.add_RZ_gate(t[0], float(rng.uniform(0, 6)))
        elif k == 4:
            c.add_gate(gate.RandomUnitary(t, int(rng.integers(1 << 30))))
        else:
            c.add_gate(gate.ParametricRY(t[0], 0.4))
    return c


@pytest.mark.parametrize("block", [1, 2, 3, 4])
def test_optimizer_preserves_the_circuit_unitary(block):
    """merge_all of the optimised circuit equals merge_all of the original
    (src/cppsim/circuit_optimizer.cpp:113-235); host-only arithmetic."""
    from qulacs_amd.circuit import QuantumCircuitOptimizer

    rng = np.random.default_rng(block)
    n = 5
    c = _random_circuit(rng, n, 40)
    want = QuantumCircuitOptimizer().merge_all(c)
    assert want.get_target_index_list() == list(range(n))
    for light in (False, True):
        c2 = c.copy()
        opt = QuantumCircuitOptimizer()
        if light:
            opt.optimize_light(c2)
        else:
            opt.optimize(c2, block)
        assert c2.get_gate_count() <= c.get_gate_count()
        if not light and block >= 2:
            assert c2.get_gate_count() < c.get_gate_count()
        for i in range(c2.get_gate_count()):
            g = c2.get_gate(i)
            if not light:
                assert len(g.get_target_index_list()) + len(g.get_control_index_list()) <= max(block, 2)
        got = QuantumCircuitOptimizer().merge_all(c2)
        assert np.allclose(got.get_matrix(), want.get_matrix(), atol=1e-12), (block, light)


def _optimized_light(c):
    from qulacs_amd.circuit import QuantumCircuitOptimizer

    c2 = c.copy()
    QuantumCircuitOptimizer().optimize_light(c2)
    merged = [QuantumCircuitOptimizer().merge_all(x).get_matrix() for x in (c2, c)]
    return c2, np.allclose(merged[0], merged[1], atol=1e-12)


def test_optimize_light_depth_and_gate_counts():
    """The fixed cases of test/cppsim/test_circuit_optimize_light.cpp:20-175 with
    their depth and gate-count expectations; the circuit unitary is unchanged."""
    n = 4
    cases = []
    c = q.QuantumCircuit(n)  # successive gates on one qubit merge
    c.add_X_gate(0)
    c.add_Y_gate(0)
    cases.append((c, 1, 1))
    c = q.QuantumCircuit(n)  # no tensor products of disjoint gates
    c.add_X_gate(0)
    c.add_Y_gate(1)
    cases.append((c, 1, 2))
    c = q.QuantumCircuit(n)  # Z on a CNOT's target does not commute through it
    c.add_Z_gate(1)
    c.add_CNOT_gate(0, 1)
    c.add_Z_gate(1)
    cases.append((c, 2, 2))
    c = q.QuantumCircuit(n)  # merged controlled gates
    c.add_Z_gate(0)
    c.add_gate(gate.merge(gate.CNOT(0, 1), gate.Y(2)))
    c.add_gate(gate.merge(gate.CNOT(1, 2), gate.Y(2)))
    c.add_Z_gate(1)
    cases.append((c, 2, 2))
    for i, (c, depth, count) in enumerate(cases):
        c2, same = _optimized_light(c)
        assert same, i
        assert (c2.calculate_depth(), c2.get_gate_count()) == (depth, count), i

    # parametric gates are left as they are (test_circuit_optimize_light.cpp:150-175)
    p = q.ParametricQuantumCircuit(2)
    p.add_X_gate(0)
    p.add_parametric_RX_gate(0, 0.1)
    p.add_parametric_RY_gate(0, 0.1)
    p2, same = _optimized_light(p)
    assert same
    assert (p2.calculate_depth(), p2.get_gate_count(), p2.get_parameter_count()) == (3, 3, 2)


@pytest.mark.parametrize("kind", ["cnot_cz", "cnot_cz_swap", "random_pairs"])
def test_optimize_light_random_circuits(kind):
    """The random families of test_circuit_optimize_light.cpp:178-315 (nearest-
    neighbour CNOT/CZ; sqrtX/sqrtY/T with any-pair CNOT/CZ/SWAP; two-qubit random
    unitaries on shuffled pairs): the optimised circuit has the same unitary."""
    rng = np.random.default_rng(len(kind))
    n = 5
    for _ in range(3):
        c = q.QuantumCircuit(n)
        if kind == "random_pairs":
            for _ in range(10 * n):
                pair = [int(x) for x in rng.permutation(n)[:2]]
                c.add_gate(gate.RandomUnitary(pair, int(rng.integers(1 << 30))))
        else:
            for _ in range(5 if kind == "cnot_cz" else 10):
                for i in range(n):
                    r = int(rng.integers(2)) + 3 if kind == "cnot_cz" else int(rng.integers(6))
                    other = int(rng.integers(n))
                    if other == i:
                        other = (other + 1) % n
                    if kind == "cnot_cz":
                        other = i + 1
                    if r == 0:
                        c.add_sqrtX_gate(i)
                    elif r == 1:
                        c.add_sqrtY_gate(i)
                    elif r == 2:
                        c.add_T_gate(i)
                    elif i + 1 < n:
                        (c.add_CNOT_gate, c.add_CZ_gate, c.add_SWAP_gate)[r - 3](i, other)
        c2, same = _optimized_light(c)
        assert same
        assert c2.get_gate_count() <= c.get_gate_count()
