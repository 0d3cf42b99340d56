"""Controlled dense gates on density matrices and state vectors, the families of
test/cppsim/test_gate_dm.cpp (MultiTarget, MultiControlSingleTarget,
SingleControlMultiTarget, MultiControlMultiTarget): random unitaries on 1-3 targets
with 0-3 controls of either value, against U rho U^dagger and U psi built in numpy."""
from __future__ import annotations

import numpy as np
import pytest

import qulacs_amd as q
from qulacs_amd import gate

pytestmark = pytest.mark.gpu
N = 6


def _full(g, n):
    m = np.asarray(g.get_matrix())
    t = list(g.get_target_index_list())
    cs = list(zip(g.get_control_index_list(), g.get_control_value_list()))
    d = 1 << n
    u = np.zeros((d, d), dtype=complex)
    for col in range(d):
        if any(((col >> c) & 1) != v for c, v in cs):
            u[col, col] = 1
            continue
        ci = sum(((col >> tq) & 1) << k for k, tq in enumerate(t))
        base = col & ~sum(1 << tq for tq in t)
        for ri in range(1 << len(t)):
            row = base | sum(((ri >> k) & 1) << tq for k, tq in enumerate(t))
            u[row, col] += m[ri, ci]
    return u


@pytest.mark.parametrize("targets,controls", [(1, 0), (2, 0), (3, 0), (1, 1), (1, 3), (2, 1), (3, 1), (2, 2), (3, 3)])
def test_controlled_dense_on_dm_and_sv(targets, controls):
    rng = np.random.default_rng(10 * targets + controls)
    for rep in range(3):
        qubits = [int(x) for x in rng.permutation(N)[: targets + controls]]
        g = gate.RandomUnitary(qubits[:targets], int(rng.integers(1 << 30)))
        for c in qubits[targets:]:
            g.add_control_qubit(c, int(rng.integers(2)))
        u = _full(g, N)
        sv = q.QuantumState(N)
        sv.set_Haar_random_state(rep + 1)
        psi = sv.get_vector()
        dm = q.DensityMatrix(N)
        dm.set_Haar_random_state(rep + 7)
        rho = dm.get_matrix()
        print("case", targets, controls, qubits, flush=True)
        g.update_quantum_state(sv)
        g.update_quantum_state(dm)
        np.testing.assert_allclose(sv.get_vector(), u @ psi, atol=1e-12)
        np.testing.assert_allclose(dm.get_matrix(), u @ rho @ u.conj().T, atol=1e-12)


@pytest.mark.parametrize("n", [6, 9])
def test_monomial_two_qubit_gates_on_dm(n):
    """Two-qubit monomial unitaries (CNOT either way, CZ, SWAP, iSWAP, a
    controlled phase with control value 0) take the one-pass path that moves
    only the density-matrix entries that change; each equals U rho U^dagger."""
    rng = np.random.default_rng(40 + n)
    iswap = np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]])
    for rep in range(4):
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        if rep == 0:
            a, b = 0, n - 1  # lowest and highest qubit: cached and streaming member accesses
        gs = [gate.CNOT(a, b), gate.CNOT(b, a), gate.CZ(a, b), gate.SWAP(a, b), gate.DenseMatrix([a, b], iswap)]
        cs = gate.DenseMatrix(b, np.diag([1, 1j]))
        cs.add_control_qubit(a, 0)
        gs.append(cs)
        for g in gs:
            dm = q.DensityMatrix(n)
            dm.set_Haar_random_state(rep + 3)
            rho = dm.get_matrix()
            u = _full(g, n)
            g.update_quantum_state(dm)
            np.testing.assert_allclose(dm.get_matrix(), u @ rho @ u.conj().T, atol=1e-12)
