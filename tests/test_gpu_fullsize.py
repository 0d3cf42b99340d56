"""Full-size checks: a 30-qubit state vector (16 GiB) and a 14-qubit density
matrix (4^14 entries, 4 GiB), the single-GPU shapes of the gate-sweep workload.

At these sizes the oracle cannot replay every gate quickly, so the checks use
properties that do not depend on size: unitarity (norm kept, a sequence
followed by its inverse restores the state), exact involutions (X, CNOT, SWAP
twice is bit-exact), closed forms (<Z> after RY), trace preservation of
channels, and sampling compared index by index with the sequential CPU loop on
the downloaded state.
"""
from __future__ import annotations

import numpy as np
import pytest

from qulacs_amd import _lib
from tests.golden.gen import haar_unitary
from tests.gpu_util import DevState, cmat, rel, u32

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N_SV = 30
N_DM = 14


def _gate_list(rng, n, count):
    gates = []
    for i in range(count):
        kind = i % 6
        if kind == 0:
            t = [int(rng.integers(n))]
        elif kind == 1:
            t = rng.choice(n, 2, replace=False).tolist()
        elif kind == 2:
            t = rng.choice(n, 5, replace=False).tolist()
        elif kind == 3:
            t = [int(x) for x in (0, n - 1)]
        elif kind == 4:
            t = [n - 1, int(rng.integers(1, 6))]
        else:
            t = [int(rng.integers(0, 3)), int(rng.integers(n - 3, n)), int(rng.integers(6, n - 3))]
        gates.append((t, haar_unitary(rng, len(t))))
    return gates


def test_30q_unitarity_inverse_and_involutions():
    rng = np.random.default_rng(30)
    dim = 1 << N_SV
    c = _lib.call
    with DevState(dim) as s, DevState(dim) as keep:
        c("initialize_Haar_random_state_with_seed_host", s.ptr, dim, 11, None, 0)
        c("copy_quantum_state_from_device_to_device", keep.ptr, s.ptr, dim, None, 0)
        gates = _gate_list(rng, N_SV, 24)
        for t, u in gates:
            print("gate", t, flush=True)
            c("multi_qubit_dense_matrix_gate_host", u32(t), len(t), cmat(u), s.ptr, dim, None, 0)
        for cc, tt in [(0, 29), (29, 3), (12, 13)]:
            c("CNOT_gate_host", cc, tt, s.ptr, dim, None, 0)
            c("CZ_gate_host", tt, cc, s.ptr, dim, None, 0)
        c("RZ_gate_host", 29, 0.37, s.ptr, dim, None, 0)
        c("multi_qubit_Pauli_rotation_gate_partial_list_host", u32([0, 15, 29]), u32([1, 2, 3]), 3, 1.1, s.ptr, dim,
          None, 0)
        nrm = c("state_norm_squared_host", s.ptr, dim, None, 0)
        assert abs(nrm - 1.0) < 1e-10
        # undo everything
        c("multi_qubit_Pauli_rotation_gate_partial_list_host", u32([0, 15, 29]), u32([1, 2, 3]), 3, -1.1, s.ptr, dim,
          None, 0)
        c("RZ_gate_host", 29, -0.37, s.ptr, dim, None, 0)
        for cc, tt in reversed([(0, 29), (29, 3), (12, 13)]):
            c("CZ_gate_host", tt, cc, s.ptr, dim, None, 0)
            c("CNOT_gate_host", cc, tt, s.ptr, dim, None, 0)
        for t, u in reversed(gates):
            c("multi_qubit_dense_matrix_gate_host", u32(t), len(t), cmat(u.conj().T), s.ptr, dim, None, 0)
        z = c("inner_product_host", keep.ptr, s.ptr, dim, None, 0)
        # |<keep|s>| = 1 and ||s - keep||^2 = 2 - 2 Re<keep|s>
        assert abs(z.re - 1.0) < 1e-10 and abs(z.im) < 1e-10
        # exact involutions: bit-identical after applying twice (compared by the inner product being exactly 1
        # would need rounding; compare checksums of the downloaded halves instead)
        c("copy_quantum_state_from_device_to_device", keep.ptr, s.ptr, dim, None, 0)
        for name, a, b in [("X", 29, None), ("CNOT", 2, 29), ("SWAP", 0, 28)]:
            for _ in range(2):
                if b is None:
                    c(f"{name}_gate_host", a, s.ptr, dim, None, 0)
                else:
                    c(f"{name}_gate_host", a, b, s.ptr, dim, None, 0)
        a = s.download()
        b = keep.download()
        assert np.array_equal(a, b)


def test_30q_closed_form_expectations():
    dim = 1 << N_SV
    c = _lib.call
    with DevState(dim) as s:
        c("initialize_quantum_state_host", s.ptr, dim, None, 0)
        c("RY_gate_host", 29, 0.81, s.ptr, dim, None, 0)
        c("RX_gate_host", 0, 1.27, s.ptr, dim, None, 0)
        ez = c("expectation_value_multi_qubit_Pauli_operator_partial_list_host", u32([29]), u32([3]), 1, s.ptr, dim,
               None, 0)
        assert abs(ez - np.cos(0.81)) < 1e-12
        # RX(theta)|0> = cos|0> + i sin|1>: <Y> = sin(theta) under exp(+i theta X / 2)
        ey = c("expectation_value_multi_qubit_Pauli_operator_partial_list_host", u32([0]), u32([2]), 1, s.ptr, dim,
               None, 0)
        assert abs(ey - np.sin(1.27)) < 1e-12
        m0 = c("M0_prob_host", 29, s.ptr, dim, None, 0)
        assert abs(m0 - np.cos(0.81 / 2) ** 2) < 1e-12


def test_30q_sampling_identical_to_sequential_cpu(oracle):
    rng = np.random.default_rng(31)
    dim = 1 << N_SV
    c = _lib.call
    with DevState(dim) as s:
        c("initialize_Haar_random_state_with_seed_host", s.ptr, dim, 5, None, 0)
        for t, u in _gate_list(rng, N_SV, 6):
            c("multi_qubit_dense_matrix_gate_host", u32(t), len(t), cmat(u), s.ptr, dim, None, 0)
        us = rng.random(1024)
        got = np.zeros(us.size, dtype=np.uint64)
        c("qm_sampling_host", us.ctypes.data_as(_lib.PD), us.size, got.ctypes.data_as(_lib.PI), s.ptr, dim, None, 0)
        host = s.download()
    oracle.set_num_threads(1)
    want = oracle.sampling(host, us)
    del host
    assert np.array_equal(got, want)


def test_14q_density_matrix_trace_hermiticity_and_pure_consistency():
    rng = np.random.default_rng(14)
    n = N_DM
    dim = 1 << n
    c = _lib.call
    with DevState(dim) as sv, DevState(dim * dim) as dm:
        c("initialize_quantum_state_host", sv.ptr, dim, None, 0)
        c("dm_initialize_quantum_state_host", dm.ptr, dim, None, 0)
        for t, u in _gate_list(rng, n, 12):
            c("multi_qubit_dense_matrix_gate_host", u32(t), len(t), cmat(u), sv.ptr, dim, None, 0)
            c("dm_multi_qubit_dense_matrix_gate_host", u32(t), len(t), cmat(u), dm.ptr, dim, None, 0)
        psi = sv.download()
        rho = dm.download().reshape(dim, dim)
        # pure evolution: rho == |psi><psi| (compared on a block of rows to bound host memory)
        rows = slice(0, 512)
        assert rel(rho[rows], np.outer(psi[rows], psi.conj())) < 1e-10
        del rho
        for q in range(0, n, 3):
            c("dm_depolarizing_host", q, 0.01, dm.ptr, dim, None, 0)
        tr = c("dm_state_norm_squared_host", dm.ptr, dim, None, 0)
        assert abs(tr - 1.0) < 1e-10
        rho = dm.download().reshape(dim, dim)
        assert np.max(np.abs(rho - rho.conj().T)) < 1e-12


def _apply_sparse(amp, targets, m, controls=()):
    """Apply matrix m (bit i of its index <-> targets[i]) to a sparse state {index: amplitude}."""
    out = {}
    for idx, a in amp.items():
        if any(not (idx >> c) & 1 for c in controls):
            out[idx] = out.get(idx, 0) + a
            continue
        col = sum(((idx >> t) & 1) << i for i, t in enumerate(targets))
        base = idx & ~sum(1 << t for t in targets)
        for row in range(1 << len(targets)):
            v = m[row, col]
            if v != 0:
                j = base | sum(((row >> i) & 1) << t for i, t in enumerate(targets))
                out[j] = out.get(j, 0) + v * a
    return {k: v for k, v in out.items() if abs(v) > 0}


def test_33q_index_math_above_32_bits():
    """Every kernel family on a 33-qubit state (128 GiB, indices above 2^32) started from
    a basis state with high bits set, against a sparse host simulation of the same gates,
    read back amplitude by amplitude."""
    import qulacs_amd as q
    from qulacs_amd import gate

    n = 33
    rng = np.random.default_rng(33)
    b0 = (1 << 32) | (1 << 31) | 3
    s = q.QuantumState(n)
    s.set_computational_basis(b0)
    amp = {b0: 1.0 + 0j}
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    H = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2)
    c, sn = np.cos(0.3), np.sin(0.3)
    steps = [
        (gate.X(32), [32], X, ()),
        (gate.CNOT(31, 32), [32], X, (31,)),
        (gate.H(0), [0], H, ()),
        (gate.RZ(32, 0.6), [32], np.diag([c + 1j * sn, c - 1j * sn]), ()),
    ]
    u2 = haar_unitary(rng, 2)
    steps.append((gate.DenseMatrix([32, 5], u2), [32, 5], u2, ()))
    u5 = haar_unitary(rng, 5)
    steps.append((gate.DenseMatrix([28, 32, 1, 30, 17], u5), [28, 32, 1, 30, 17], u5, ()))
    u3 = haar_unitary(rng, 3)
    steps.append((gate.DenseMatrix([32, 2, 20], u3), [32, 2, 20], u3, ()))
    pz = np.kron(np.diag([1.0, -1.0]), np.array([[0, 1], [1, 0]]))  # X on qubit 4, Z on qubit 32
    th = 0.45
    steps.append((gate.PauliRotation([4, 32], [1, 3], th), [4, 32],
                  np.cos(th / 2) * np.eye(4) + 1j * np.sin(th / 2) * pz, ()))
    steps.append((gate.CZ(32, 31), [31], np.diag([1.0, -1.0]).astype(complex), (32,)))
    # qubits below 3 with a partial operation or a control there: the line kernel
    steps.append((gate.CZ(1, 32), [32], np.diag([1.0, -1.0]).astype(complex), (1,)))
    steps.append((gate.CNOT(2, 31), [31], X, (2,)))
    steps.append((gate.S(0), [0], np.diag([1.0, 1j]), ()))
    steps.append((gate.CZ(0, 1), [1], np.diag([1.0, -1.0]).astype(complex), (0,)))
    for g, t, m, ctl in steps:
        print("gate", g.get_name(), t, flush=True)
        g.update_quantum_state(s)
        amp = _apply_sparse(amp, t, m, ctl)
    assert len(amp) <= 1024
    for idx, a in amp.items():
        assert abs(s.get_amplitude(idx) - a) < 1e-12, (idx, s.get_amplitude(idx), a)
    assert abs(s.get_squared_norm() - 1.0) < 1e-12
    p32 = sum(abs(a) ** 2 for k, a in amp.items() if not (k >> 32) & 1)
    assert abs(s.get_zero_probability(32) - p32) < 1e-12
    obs = q.Observable(n)
    obs.add_operator(0.5, "Z 32 Z 0")
    want = 0.5 * sum(abs(a) ** 2 * (1 - 2 * ((k >> 32) & 1)) * (1 - 2 * (k & 1)) for k, a in amp.items())
    assert abs(obs.get_expectation_value(s) - want) < 1e-12


def test_16q_density_matrix_cdf_on_the_multi_block_path():
    """A 16-qubit density matrix (4^16 entries, 64 GiB) has a 2^16-entry diagonal,
    the size from which the CDF runs on the multi-block chunk walk.  Built on the
    device from a real pure state, so every diagonal entry is real (|rho_ii| is
    the entry itself); the diagonal is read back entry by entry and the device CDF
    must equal its left-to-right sums bit for bit, with samples to match."""
    from itertools import accumulate

    n = 16
    dim = 1 << n
    rng = np.random.default_rng(16)
    psi = rng.standard_normal(dim) * np.exp(-np.arange(dim) / (dim / 7))  # real, decaying: many binades
    psi /= np.linalg.norm(psi)
    c = _lib.call
    with DevState.from_host(psi.astype(np.complex128)) as sv, DevState(dim * dim) as dm:
        c("dm_initialize_with_pure_state_host", dm.ptr, sv.ptr, dim, None, 0)
        diag = np.zeros(dim, dtype=np.complex128)
        one = np.zeros(1, dtype=np.complex128)
        for i in range(dim):
            c("copy_quantum_state_from_device_to_host", one.ctypes.data, dm.ptr + 16 * i * (dim + 1), 1, None, 0)
            diag[i] = one[0]
        assert np.all(diag.imag == 0) and np.all(diag.real >= 0)
        want = np.array([0.0] + list(accumulate(diag.real.tolist())))
        got = np.zeros(dim + 1)
        c("dm_sampling_cdf_host", got.ctypes.data_as(_lib.PD), dm.ptr, dim, None, 0)
        bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
        assert bad.size == 0, bad[:5]
        us = np.concatenate([rng.random(2000) * want[-1], want[1:2000:7], np.nextafter(want[1:2000:7], 2.0)])
        out = np.zeros(us.size, dtype=np.uint64)
        c("dm_sampling_host", us.ctypes.data_as(_lib.PD), us.size, out.ctypes.data_as(_lib.PI), dm.ptr, dim, None, 0)
        expect = (np.searchsorted(want, us, side="left") - 1).astype(np.int64).astype(np.uint64)
        assert np.array_equal(out, expect)
