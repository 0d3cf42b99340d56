"""Density-matrix parity on the GPU, through the C ABI (dm_*_host)."""
from __future__ import annotations

import numpy as np
import pytest

from qulacs_amd import _lib
from tests.golden.gen import PAULI, expand, haar_unitary, pauli_matrix, random_pauli_terms, random_state
from tests.gpu_util import DevState, cmat, rel, u32
from tests.test_oracle import NAMED_MATRICES, rot_matrix

pytestmark = pytest.mark.gpu


def _mixed(oracle, rng, n, k=3):
    rho = np.zeros(4**n, dtype=np.complex128)
    for w in rng.dirichlet(np.ones(k)):
        rho += w * oracle.dm_from_pure(random_state(rng, n))
    return rho


def test_golden_dm_through_c_abi(golden):
    g = golden("dm.npz")
    n = int(g["n"])
    dim = 1 << n
    c = _lib.call
    with DevState(dim * dim) as s:
        for i in range(len(g["name"])):
            name = str(g["name"][i])
            tg = [int(x) for x in g["targets"][i] if x >= 0]
            cl = [int(x) for x in g["controls"][i] if x >= 0]
            cv = [int(x) for x in g["cvals"][i] if x >= 0]
            k = len(tg)
            mat = g["matrix"][i][: 1 << k, : 1 << k]
            s.upload(g["rho_in"][i])
            if name == "dm_dense1":
                c("dm_single_qubit_dense_matrix_gate_host", tg[0], cmat(mat), s.ptr, dim, None, 0)
            elif name == "dm_dense":
                c("dm_multi_qubit_dense_matrix_gate_host", u32(tg), k, cmat(mat), s.ptr, dim, None, 0)
            elif name == "dm_cdense":
                c("dm_multi_qubit_control_multi_qubit_dense_matrix_gate_host", u32(cl), u32(cv), len(cl), u32(tg), k,
                  cmat(mat), s.ptr, dim, None, 0)
            elif name in ("dm_CNOT", "dm_CZ"):
                m = PAULI[1] if name == "dm_CNOT" else PAULI[3]
                c("dm_multi_qubit_control_multi_qubit_dense_matrix_gate_host", u32(cl), u32([1]), 1, u32(tg), 1,
                  cmat(m), s.ptr, dim, None, 0)
            elif name == "dm_SWAP":
                c("dm_multi_qubit_dense_matrix_gate_host", u32([cl[0], tg[0]]), 2, cmat(np.eye(4)[[0, 2, 1, 3]]),
                  s.ptr, dim, None, 0)
            elif name in ("dm_RX", "dm_RY", "dm_RZ"):
                c("dm_single_qubit_dense_matrix_gate_host", tg[0], cmat(rot_matrix(name[3:], float(g["param"][i]))),
                  s.ptr, dim, None, 0)
            else:
                c("dm_single_qubit_dense_matrix_gate_host", tg[0], cmat(NAMED_MATRICES[name[3:]]), s.ptr, dim, None, 0)
            assert rel(s.download(), g["rho_out"][i]) < 1e-12, name
        s.upload(g["rho"])
        assert abs(c("dm_state_norm_squared_host", s.ptr, dim, None, 0) - g["trace"]) < 1e-12
        for t in range(n):
            assert abs(c("dm_M0_prob_host", t, s.ptr, dim, None, 0) - g["M0"][t]) < 1e-12
        assert abs(c("dm_marginal_prob_host", u32([0, 2]), u32([1, 0]), 2, s.ptr, dim, None, 0) - g["marginal"]) < 1e-12
        for k in range(len(g["expectation"])):
            t = [int(x) for x in g["term_targets"][k] if x >= 0]
            p = [int(x) for x in g["term_paulis"][k] if x >= 0]
            e = c("dm_expectation_value_multi_qubit_Pauli_operator_partial_list_host", u32(t), u32(p), len(t), s.ptr,
                  dim, None, 0)
            assert abs(e - g["expectation"][k]) < 1e-12


@pytest.mark.parametrize("n", [5, 8])
def test_dm_random_sequence_matches_oracle(oracle, n):
    rng = np.random.default_rng(300 + n)
    dim = 1 << n
    rho = _mixed(oracle, rng, n)
    ref = rho.copy()
    c = _lib.call
    with DevState.from_host(rho) as s:
        for step in range(24):
            kind = step % 6
            if kind == 0:
                t = int(rng.integers(n)); u = haar_unitary(rng, 1)
                oracle.dm_dense1(t, u, ref, dim)
                c("dm_single_qubit_dense_matrix_gate_host", t, cmat(u), s.ptr, dim, None, 0)
            elif kind in (1, 2):
                k = 2 if kind == 1 else 3
                tg = rng.choice(n, k, replace=False).tolist(); u = haar_unitary(rng, k)
                oracle.dm_dense(tg, u, ref, dim)
                c("dm_multi_qubit_dense_matrix_gate_host", u32(tg), k, cmat(u), s.ptr, dim, None, 0)
            elif kind == 3:
                q = rng.choice(n, 3, replace=False).tolist(); u = haar_unitary(rng, 1)
                oracle.dm_controlled_dense(q[:2], [1, 0], q[2:], u, ref, dim)
                c("dm_multi_qubit_control_multi_qubit_dense_matrix_gate_host", u32(q[:2]), u32([1, 0]), 2, u32(q[2:]),
                  1, cmat(u), s.ptr, dim, None, 0)
            elif kind == 4:
                t = int(rng.integers(n)); p = float(rng.uniform(0, 0.3))
                oracle.dm_depolarizing(t, p, ref, dim)
                c("dm_depolarizing_host", t, p, s.ptr, dim, None, 0)
            else:
                if n >= 5:
                    tg = rng.choice(n, 5, replace=False).tolist(); u = haar_unitary(rng, 5)
                    oracle.dm_dense(tg, u, ref, dim)
                    c("dm_multi_qubit_dense_matrix_gate_host", u32(tg), 5, cmat(u), s.ptr, dim, None, 0)
        out = s.download()
        assert rel(out, ref) < 1e-10
        for tg, pl, _ in random_pauli_terms(rng, n, 6):
            e = c("dm_expectation_value_multi_qubit_Pauli_operator_partial_list_host", u32(tg), u32(pl), len(tg),
                  s.ptr, dim, None, 0)
            assert abs(e - oracle.dm_expectation(tg, pl, ref, dim)) < 1e-11
        assert abs(c("dm_state_norm_squared_host", s.ptr, dim, None, 0) - oracle.dm_norm(ref, dim)) < 1e-11
        us = rng.random(500)
        got = np.zeros(us.size, dtype=np.uint64)
        c("dm_sampling_host", us.ctypes.data_as(_lib.PD), us.size, got.ctypes.data_as(_lib.PI), s.ptr, dim, None, 0)
        # the samples follow the device matrix's own diagonal
        assert np.array_equal(got, oracle.dm_sampling(out, dim, us))


def test_superoperator_general_kraus(oracle):
    """Amplitude damping as a superoperator equals sum_i K_i rho K_i^dagger."""
    rng = np.random.default_rng(9)
    n = 4
    dim = 1 << n
    g = 0.3
    K = [np.array([[1, 0], [0, np.sqrt(1 - g)]]), np.array([[0, np.sqrt(g)], [0, 0]])]
    # member index = row_bit * 2 + column_bit: S[(oy,ox),(iy,ix)] = sum K[oy,iy] conj(K[ox,ix])
    S = sum(np.kron(k, k.conj()) for k in K)
    rho = _mixed(oracle, rng, n)
    R = rho.reshape(dim, dim)
    for t in (0, 3):
        Ks = [expand(n, [t], k) for k in K]
        want = sum(k @ R @ k.conj().T for k in Ks)
        with DevState.from_host(rho) as s:
            _lib.call("dm_superoperator_host", u32([t]), 1, cmat(S), s.ptr, dim, None, 0)
            assert rel(s.download().reshape(dim, dim), want) < 1e-12
    # two-qubit superoperator: K (x) K on (1, 2)
    K2 = [np.kron(a, b) for a in K for b in K]
    S2 = sum(np.kron(k, k.conj()) for k in K2)
    Ks = [expand(n, [1, 2], k) for k in K2]
    want = sum(k @ R @ k.conj().T for k in Ks)
    with DevState.from_host(rho) as s:
        _lib.call("dm_superoperator_host", u32([1, 2]), 2, cmat(S2), s.ptr, dim, None, 0)
        assert rel(s.download().reshape(dim, dim), want) < 1e-12


def test_dm_pure_state_matches_state_vector(oracle):
    rng = np.random.default_rng(10)
    n = 6
    dim = 1 << n
    psi = random_state(rng, n)
    c = _lib.call
    with DevState.from_host(psi) as sv, DevState(dim * dim) as dm:
        c("dm_initialize_with_pure_state_host", dm.ptr, sv.ptr, dim, None, 0)
        for _ in range(10):
            tg = rng.choice(n, 2, replace=False).tolist(); u = haar_unitary(rng, 2)
            c("multi_qubit_dense_matrix_gate_host", u32(tg), 2, cmat(u), sv.ptr, dim, None, 0)
            c("dm_multi_qubit_dense_matrix_gate_host", u32(tg), 2, cmat(u), dm.ptr, dim, None, 0)
        v = sv.download()
        assert rel(dm.download().reshape(dim, dim), np.outer(v, v.conj())) < 1e-12
        c("dm_initialize_quantum_state_host", dm.ptr, dim, None, 0)
        e = np.zeros(dim * dim, dtype=np.complex128); e[0] = 1
        assert np.array_equal(dm.download(), e)
