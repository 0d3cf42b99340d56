"""The qulacs Python API (qulacs_amd) on the GPU, mirroring the reference's
python/tests/single_cpu tests plus parity of every gate factory, noise channel,
observable, sampling call and gradient against the oracle or closed forms.

Bar: states within 1e-10 relative L2; scalars within 1e-10 (absolute, unit-norm
states); sampled indices identical for identical uniforms; gradients from the
adjoint sweep within 1e-9 of the parameter-shift rule and 1e-6 of central
finite differences (step 1e-5).
"""
from __future__ import annotations

import numpy as np
import pytest

import qulacs_amd as q
from qulacs_amd import gate
from tests.gpu_util import rel

pytestmark = pytest.mark.gpu

TOL = 1e-10


def full_operator(m, targets, n, controls=(), cvals=()):
    """Dense 2^n operator of matrix m on `targets` (bit i <-> targets[i]) with controls."""
    dim = 1 << n
    out = np.zeros((dim, dim), dtype=complex)
    k = len(targets)
    for col in range(dim):
        if any(((col >> c) & 1) != v for c, v in zip(controls, cvals)):
            out[col, col] = 1
            continue
        x = sum(((col >> t) & 1) << i for i, t in enumerate(targets))
        base = col & ~sum(1 << t for t in targets)
        for y in range(1 << k):
            row = base | sum(((y >> i) & 1) << t for i, t in enumerate(targets))
            out[row, col] += m[y, x]
    return out


def gate_operator(g, n):
    return full_operator(g.get_matrix(), g.get_target_index_list(), n, g.get_control_index_list(),
                         g.get_control_value_list())


def random_gate_list(rng, n):
    """One of every factory on random qubits."""
    def qs(k):
        return [int(x) for x in rng.choice(n, k, replace=False)]

    a = lambda: float(rng.uniform(0, 2 * np.pi))  # noqa: E731
    t = lambda: int(rng.integers(n))  # noqa: E731
    gs = [
        gate.X(t()), gate.Y(t()), gate.Z(t()), gate.H(t()), gate.S(t()), gate.Sdag(t()), gate.T(t()),
        gate.Tdag(t()), gate.sqrtX(t()), gate.sqrtXdag(t()), gate.sqrtY(t()), gate.sqrtYdag(t()),
        gate.RX(t(), a()), gate.RY(t(), a()), gate.RZ(t(), a()), gate.RotX(t(), a()), gate.RotY(t(), a()),
        gate.RotZ(t(), a()), gate.U1(t(), a()), gate.U2(t(), a(), a()), gate.U3(t(), a(), a(), a()),
        gate.CNOT(*qs(2)), gate.CZ(*qs(2)), gate.SWAP(*qs(2)),
        gate.RandomUnitary(qs(1), 11), gate.RandomUnitary(qs(2), 12), gate.RandomUnitary(qs(3), 13),
        gate.RandomUnitary(qs(4), 14), gate.RandomUnitary(qs(5), 15),
        gate.DiagonalMatrix(qs(2), np.exp(1j * rng.uniform(0, 6, 4))),
        gate.Pauli(qs(3), [1, 2, 3]), gate.PauliRotation(qs(3), [3, 1, 2], a()),
        gate.ParametricRX(t(), a()), gate.ParametricRY(t(), a()), gate.ParametricRZ(t(), a()),
        gate.ParametricPauliRotation(qs(2), [2, 2], a()),
        gate.merge(gate.H(0), gate.RY(1, 0.3)),
    ]
    g = gate.to_matrix_gate(gate.RandomUnitary(qs(1), 16))
    c = [x for x in range(n) if x not in g.get_target_index_list()]
    g.add_control_qubit(c[0], 1)
    g.add_control_qubit(c[1], 0)
    gs.append(g)
    g2 = gate.DenseMatrix(qs(2), gate.RandomUnitary([0, 1], 17).get_matrix())
    c = [x for x in range(n) if x not in g2.get_target_index_list()]
    g2.add_control_qubit(c[0], 1)
    gs.append(g2)
    order = rng.permutation(len(gs))
    return [gs[i] for i in order]


# ------------------------------------------------------------------ reference python tests
def test_state_basics():
    # python/tests/single_cpu/test_state.py
    n, dim = 4, 16
    s = q.QuantumState(n)
    assert s.get_device_name() == "gpu" and s.get_qubit_count() == n
    v = s.get_vector()
    assert len(v) == dim and v[0] == 1 and np.count_nonzero(v) == 1
    s.set_computational_basis(0b0101)
    v = s.get_vector()
    assert v[0b0101] == 1 and np.count_nonzero(v) == 1
    assert s.get_amplitude(0b0101) == 1
    s.set_Haar_random_state(7)
    assert abs(s.get_squared_norm() - 1) < TOL
    c = s.copy()
    assert np.array_equal(c.get_vector(), s.get_vector())
    host = np.arange(dim) + 1j
    s.load(host)
    assert np.array_equal(s.get_vector(), host)
    s.load(list(host))
    assert np.array_equal(s.get_vector(), host)
    with pytest.raises(RuntimeError):
        s.load(np.zeros(dim + 1))
    with pytest.raises(RuntimeError):
        s.set_computational_basis(dim)


def test_bell_state():
    # python/tests/single_cpu/test_circuit.py
    s = q.QuantumState(4)
    c = q.QuantumCircuit(4)
    c.add_H_gate(0)
    c.add_CNOT_gate(0, 1)
    c.update_quantum_state(s)
    want = np.zeros(16)
    want[0] = want[3] = np.sqrt(0.5)
    assert np.abs(s.get_vector() - want).max() < TOL


def test_density_matrix_from_pure():
    # python/tests/single_cpu/test_density_matrix_handling.py
    sv = q.QuantumState(5)
    dm = q.DensityMatrix(5)
    sv.set_Haar_random_state(0)
    dm.load(sv)
    v = sv.get_vector()
    assert np.abs(dm.get_matrix() - np.outer(v, v.conj())).max() < TOL
    dm2 = q.DensityMatrix(5)
    dm2.load(v)
    assert np.abs(dm2.get_matrix() - dm.get_matrix()).max() < TOL
    assert abs(dm.get_squared_norm() - 1) < TOL


# ------------------------------------------------------------------ gate factories
@pytest.mark.parametrize("n,seed", [(6, 1), (9, 2)])
def test_every_factory_on_state_vector(oracle, n, seed):
    rng = np.random.default_rng(seed)
    s = q.QuantumState(n)
    s.set_Haar_random_state(seed)
    psi = s.get_vector()
    circuit = q.QuantumCircuit(n)
    for g in random_gate_list(rng, n):
        circuit.add_gate(g)
    for i in range(circuit.get_gate_count()):
        g = circuit.get_gate(i)
        print("gate", i, g.get_name(), g.get_target_index_list(), g.get_control_index_list(), flush=True)
        g.update_quantum_state(s)
        psi = gate_operator(g, n) @ psi
        assert rel(s.get_vector(), psi) < TOL, (i, g.get_name())
    # the circuit path and the inverse circuit
    s2 = q.QuantumState(n)
    s2.set_Haar_random_state(seed)
    start = s2.get_vector()
    circuit.update_quantum_state(s2)
    assert rel(s2.get_vector(), psi) < TOL
    circuit.get_inverse().update_quantum_state(s2)
    assert rel(s2.get_vector(), start) < 1e-9


@pytest.mark.parametrize("n", [5, 6])
def test_every_unitary_factory_on_density_matrix(n):
    rng = np.random.default_rng(30 + n)
    sv = q.QuantumState(n)
    sv.set_Haar_random_state(5)
    dm = q.DensityMatrix(n)
    dm.load(sv)
    rho = dm.get_matrix()
    for g in random_gate_list(rng, n):
        print("gate", g.get_name(), g.get_target_index_list(), g.get_control_index_list(), flush=True)
        u = gate_operator(g, n)
        g.update_quantum_state(dm)
        rho = u @ rho @ u.conj().T
        assert rel(dm.get_matrix().ravel(), rho.ravel()) < TOL, g.get_name()


def test_gate_target_out_of_range_raises():
    s = q.QuantumState(3)
    with pytest.raises(RuntimeError):
        gate.X(3).update_quantum_state(s)
    with pytest.raises(RuntimeError):
        gate.CNOT(4, 0).update_quantum_state(s)
    c = q.QuantumCircuit(4)
    with pytest.raises(RuntimeError):
        c.update_quantum_state(s)


# ------------------------------------------------------------------ noise
def channel_apply(rho, weighted_kraus, targets, n):
    out = np.zeros_like(rho)
    for w, k in weighted_kraus:
        K = full_operator(k, targets, n)
        out += w * K @ rho @ K.conj().T
    return out


@pytest.mark.parametrize("p", [0.0, 0.1, 0.37])
def test_noise_channels_on_density_matrix(oracle, p):
    n = 4
    I2, X, Z = np.eye(2), np.array([[0, 1], [1, 0]]), np.diag([1, -1])
    Y = np.array([[0, -1j], [1j, 0]])
    sv = q.QuantumState(n)
    sv.set_Haar_random_state(3)
    dm = q.DensityMatrix(n)
    dm.load(sv)
    rho = dm.get_matrix()
    p1, p2 = p * (1 - p), p * p
    k0 = np.array([[1, 0], [0, np.sqrt(1 - p)]])
    k1 = np.array([[0, np.sqrt(p)], [0, 0]])
    steps = [
        (gate.DepolarizingNoise(2, p), None),
        (gate.BitFlipNoise(0, p), [(1 - p, I2), (p, X)]),
        (gate.DephasingNoise(3, p), [(1 - p, I2), (p, Z)]),
        (gate.IndependentXZNoise(1, p), [(1 - 2 * p1 - p2, I2), (p1, X), (p1, Z), (p2, Y)]),
        (gate.AmplitudeDampingNoise(2, p), [(1, k0), (1, k1)]),
        (gate.TwoQubitDepolarizingNoise(3, 0, p), "two"),
        (gate.Probabilistic([0.25, 0.5], [gate.H(1), gate.CNOT(1, 3)]), "mix"),
    ]
    for g, kraus in steps:
        print("channel", g.get_name(), g.get_target_index_list(), flush=True)
        g.update_quantum_state(dm)
        if kraus is None:
            flat = rho.ravel().copy()
            oracle.dm_depolarizing(2, p, flat, 1 << n)
            rho = flat.reshape(rho.shape)
        elif kraus == "two":
            P = [I2, X, Y, Z]
            terms = [(1 - p if i == 0 else p / 15, np.kron(P[i // 4], P[i % 4])) for i in range(16)]
            rho = channel_apply(rho, terms, [3, 0], n)
        elif kraus == "mix":
            H = gate.H(1).get_matrix()
            out = 0.25 * full_operator(H, [1], n) @ rho @ full_operator(H, [1], n).conj().T
            cx = full_operator(np.array([[0, 1], [1, 0]]), [3], n, [1], [1])
            out += 0.5 * cx @ rho @ cx.conj().T + 0.25 * rho
            rho = out
        else:
            rho = channel_apply(rho, kraus, g.get_target_index_list(), n)
        assert rel(dm.get_matrix().ravel(), rho.ravel()) < TOL, g.get_name()
    assert abs(dm.get_squared_norm() - 1) < 1e-12


def test_probabilistic_on_state_vector_follows_seeded_uniform(oracle):
    n = 3
    for seed in range(6):
        g = gate.Probabilistic([0.3, 0.3, 0.2], [gate.X(0), gate.Y(1), gate.Z(2)])
        g.set_seed(seed)
        r = oracle.mt64_uniforms(seed, 1)[0]
        branch = int(np.searchsorted(np.cumsum([0, 0.3, 0.3, 0.2]), r, side="right")) - 1
        s = q.QuantumState(n)
        s.set_Haar_random_state(1)
        psi = s.get_vector()
        g.update_quantum_state(s)
        if branch < 3:
            psi = gate_operator([gate.X(0), gate.Y(1), gate.Z(2)][branch], n) @ psi
        assert rel(s.get_vector(), psi) < TOL, (seed, branch)


def test_cptp_on_state_vector_follows_seeded_uniform(oracle):
    n, p = 3, 0.4
    for seed in range(6):
        g = gate.AmplitudeDampingNoise(1, p)
        g.set_seed(seed)
        r = oracle.mt64_uniforms(seed, 1)[0]
        s = q.QuantumState(n)
        s.set_Haar_random_state(2)
        psi = s.get_vector()
        k0 = full_operator(np.array([[1, 0], [0, np.sqrt(1 - p)]]), [1], n)
        k1 = full_operator(np.array([[0, np.sqrt(p)], [0, 0]]), [1], n)
        a = k0 @ psi
        w0 = np.vdot(a, a).real
        want = a / np.sqrt(w0) if r < w0 else (k1 @ psi) / np.linalg.norm(k1 @ psi)
        g.update_quantum_state(s)
        assert rel(s.get_vector(), want) < TOL, seed


# ------------------------------------------------------------------ statistics / observables
def test_state_statistics():
    n = 8
    rng = np.random.default_rng(9)
    s = q.QuantumState(n)
    s.set_Haar_random_state(4)
    v = s.get_vector()
    pr = np.abs(v) ** 2
    idx = np.arange(1 << n)
    for t in range(n):
        assert abs(s.get_zero_probability(t) - pr[((idx >> t) & 1) == 0].sum()) < TOL
    mv = [int(x) for x in rng.integers(0, 3, n)]
    mask = np.ones(1 << n, bool)
    for t, m in enumerate(mv):
        if m != 2:
            mask &= ((idx >> t) & 1) == m
    assert abs(s.get_marginal_probability(mv) - pr[mask].sum()) < TOL
    assert abs(s.get_entropy() - (-(pr[pr > 1e-15] * np.log(pr[pr > 1e-15])).sum())) < 1e-9
    dm = q.DensityMatrix(n)
    dm.load(s)
    assert abs(dm.get_zero_probability(3) - s.get_zero_probability(3)) < TOL
    assert abs(dm.get_marginal_probability(mv) - s.get_marginal_probability(mv)) < TOL
    assert abs(dm.get_entropy() - s.get_entropy()) < 1e-9
    # inner product / add / multiply
    s2 = q.QuantumState(n)
    s2.set_Haar_random_state(5)
    w = s2.get_vector()
    assert abs(q.state.inner_product(s, s2) - np.vdot(v, w)) < TOL
    s2.add_state_with_coef(0.5 - 0.25j, s)
    s2.multiply_coef(2j)
    assert rel(s2.get_vector(), 2j * (w + (0.5 - 0.25j) * v)) < TOL


@pytest.mark.parametrize("n", [4, 10])
def test_observable_expectation(n):
    rng = np.random.default_rng(n)
    obs = q.Observable(n)
    gen = q.GeneralQuantumOperator(n)
    names = "IXYZ"
    for _ in range(12):
        qs = rng.choice(n, int(rng.integers(1, min(n, 5) + 1)), replace=False)
        s = " ".join(f"{names[int(rng.integers(4))]} {int(t)}" for t in qs)
        obs.add_operator(float(rng.normal()), s)
        gen.add_operator(complex(rng.normal(), rng.normal()), s)
    st = q.QuantumState(n)
    st.set_Haar_random_state(8)
    v = st.get_vector()
    M, G = obs.get_matrix().toarray(), gen.get_matrix().toarray()
    e = obs.get_expectation_value(st)
    assert isinstance(e, float) and abs(e - np.vdot(v, M @ v).real) < TOL
    assert abs(gen.get_expectation_value(st) - np.vdot(v, G @ v)) < TOL
    st2 = q.QuantumState(n)
    st2.set_Haar_random_state(9)
    w = st2.get_vector()
    assert abs(gen.get_transition_amplitude(st2, st) - np.vdot(w, G @ v)) < TOL
    # apply_to_state
    work, dst = q.QuantumState(n), q.QuantumState(n)
    gen.apply_to_state(work, st, dst)
    assert rel(dst.get_vector(), G @ v) < TOL
    # the same observable on the density matrix of a mixed state
    dm = q.DensityMatrix(n)
    dm.load(st)
    gate.DepolarizingNoise(1, 0.2).update_quantum_state(dm)
    rho = dm.get_matrix()
    assert abs(obs.get_expectation_value(dm) - np.trace(M @ rho).real) < TOL
    assert abs(gen.get_expectation_value(dm) - np.trace(G @ rho)) < TOL


# ------------------------------------------------------------------ sampling
@pytest.mark.parametrize("n", [5, 14])
def test_seeded_sampling_matches_oracle(oracle, n):
    s = q.QuantumState(n)
    s.set_Haar_random_state(21)
    v = s.get_vector()
    for seed in (0, 1, 123456789):
        got = np.array(s.sampling(4096, seed), dtype=np.uint64)
        want = oracle.sampling(v, oracle.mt64_uniforms(seed, 4096))
        assert np.array_equal(got, want), seed
    dm = q.DensityMatrix(min(n, 6))
    sv = q.QuantumState(min(n, 6))
    sv.set_Haar_random_state(22)
    dm.load(sv)
    gate.DepolarizingNoise(0, 0.3).update_quantum_state(dm)
    rho = dm.get_matrix()
    got = np.array(dm.sampling(2048, 77), dtype=np.uint64)
    want = oracle.dm_sampling(rho.ravel(), rho.shape[0], oracle.mt64_uniforms(77, 2048))
    assert np.array_equal(got, want)


# ------------------------------------------------------------------ gradients
def build_vqe_circuit(n, layers, rng):
    c = q.ParametricQuantumCircuit(n)
    for _ in range(layers):
        for t in range(n):
            c.add_parametric_RX_gate(t, float(rng.uniform(0, 6)))
            c.add_parametric_RZ_gate(t, float(rng.uniform(0, 6)))
        for t in range(n - 1):
            c.add_CNOT_gate(t, t + 1)
        c.add_parametric_multi_Pauli_rotation_gate([0, n - 1], [2, 1], float(rng.uniform(0, 6)))
        c.add_parametric_RY_gate(1, float(rng.uniform(0, 6)))
        c.add_gate(gate.RandomUnitary([0, 2], int(rng.integers(1 << 30))))
    return c


def expectation(c, obs):
    s = q.QuantumState(c.get_qubit_count())
    c.update_quantum_state(s)
    return obs.get_expectation_value(s)


def test_backprop_matches_parameter_shift_and_finite_differences():
    n = 5
    rng = np.random.default_rng(44)
    c = build_vqe_circuit(n, 2, rng)
    obs = q.Observable(n)
    obs.add_operator(0.7, "Z 0 Z 1")
    obs.add_operator(-0.3, "X 2 Y 4")
    obs.add_operator(1.1, "Z 3")
    grad = np.array(c.backprop(obs))
    shift = np.array(q.GradCalculator().calculate_grad(c, obs))
    assert np.abs(shift.imag).max() < 1e-12
    assert np.abs(grad - shift.real).max() < 1e-9
    h = 1e-5
    for i in range(c.get_parameter_count()):
        th = c.get_parameter(i)
        c.set_parameter(i, th + h)
        ep = expectation(c, obs)
        c.set_parameter(i, th - h)
        em = expectation(c, obs)
        c.set_parameter(i, th)
        assert abs((ep - em) / (2 * h) - grad[i]) < 1e-6, i


# ------------------------------------------------------------------ optimizer / simulator
@pytest.mark.parametrize("block", [2, 3, 5])
def test_optimized_circuit_gives_the_same_state(block):
    from qulacs_amd.circuit import QuantumCircuitOptimizer

    n = 8
    rng = np.random.default_rng(block)
    c = q.QuantumCircuit(n)
    for _ in range(60):
        t = [int(x) for x in rng.choice(n, 2, replace=False)]
        k = int(rng.integers(5))
        if k == 0:
            c.add_gate(gate.RandomUnitary([t[0]], int(rng.integers(1 << 30))))
        elif k == 1:
            c.add_CNOT_gate(*t)
        elif k == 2:
            c.add_CZ_gate(*t)
        elif k == 3:
            c.add_gate(gate.ParametricRX(t[0], 0.3))
        else:
            c.add_RZ_gate(t[0], float(rng.uniform(0, 6)))
    s0 = q.QuantumState(n)
    s0.set_Haar_random_state(3)
    s1 = s0.copy()
    c.update_quantum_state(s0)
    c2 = c.copy()
    QuantumCircuitOptimizer().optimize(c2, block)
    assert c2.get_gate_count() < c.get_gate_count()
    c2.update_quantum_state(s1)
    assert rel(s1.get_vector(), s0.get_vector()) < TOL
    s2 = q.QuantumState(n)
    s2.set_Haar_random_state(3)
    QuantumCircuitOptimizer().merge_all(c).update_quantum_state(s2)
    assert rel(s2.get_vector(), s0.get_vector()) < TOL


def test_circuit_simulator():
    n = 6
    c = q.QuantumCircuit(n)
    for i in range(n):
        c.add_H_gate(i)
    c.add_CNOT_gate(0, 5)
    st = q.QuantumState(n)
    sim = q.QuantumCircuitSimulator(c, st)
    sim.initialize_state(0)
    sim.simulate()
    obs = q.Observable(n)
    obs.add_operator(1.0, "X 0 X 5")
    assert abs(sim.get_expectation_value(obs) - 1.0) < TOL
    sim.copy_state_to_buffer()
    sim.initialize_state(3)
    sim.swap_state_and_buffer()
    assert abs(st.get_vector()[0] - 1 / 8) < TOL
    sim.copy_state_from_buffer()
    assert st.get_vector()[3] == 1
    assert sim.get_gate_count() == n + 1
