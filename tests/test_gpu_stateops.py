"""State utilities on the device (state.tensor_product, permutate_qubit,
drop_qubit, partial_trace, make_superposition, make_mixture; reference
src/cppsim/state.cpp:32-100 and state_dm.cpp:9-90) against numpy's own
formulations: Kronecker products, axis transposes, fixed-index slices and
traces of the qubit tensor (qubit q <-> C-order axis n-1-q)."""
from __future__ import annotations

import numpy as np
import pytest

import qulacs_amd as q
from qulacs_amd import state

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _sv(n, seed):
    s = q.QuantumState(n)
    s.set_Haar_random_state(seed)
    return s


def _dm(n, seed):
    r = q.DensityMatrix(n)
    a, b = _sv(n, seed), _sv(n, seed + 1)
    m = 0.3 * np.outer(a.get_vector(), a.get_vector().conj()) + 0.7 * np.outer(b.get_vector(), b.get_vector().conj())
    r.load(m)
    return r


def test_tensor_products():
    a, b = _sv(3, 1), _sv(2, 2)
    t = state.tensor_product(a, b)
    assert t.get_qubit_count() == 5
    assert np.allclose(t.get_vector(), np.kron(a.get_vector(), b.get_vector()), atol=TOL)
    A, B = _dm(2, 3), _dm(2, 5)
    T = state.tensor_product(A, B)
    assert np.allclose(T.get_matrix(), np.kron(A.get_matrix(), B.get_matrix()), atol=TOL)


@pytest.mark.parametrize("n", [3, 6])
def test_permutate_qubit(n):
    rng = np.random.default_rng(n)
    order = [int(x) for x in rng.permutation(n)]
    s = _sv(n, 10 + n)
    p = state.permutate_qubit(s, order)
    # destination qubit k holds source qubit order[k]
    axes = [n - 1 - order[n - 1 - a] for a in range(n)]
    want = np.transpose(s.get_vector().reshape((2,) * n), axes).reshape(-1)
    assert np.allclose(p.get_vector(), want, atol=0)
    if n == 3:
        r = _dm(n, 7)
        pr = state.permutate_qubit(r, order)
        t = r.get_matrix().reshape((2,) * (2 * n))
        want = np.transpose(t, axes + [n + a for a in axes]).reshape(1 << n, 1 << n)
        assert np.allclose(pr.get_matrix(), want, atol=0)
    # a 2-cycle is the SWAP gate
    sw = _sv(3, 4)
    got = state.permutate_qubit(sw, [1, 0, 2]).get_vector()
    q.gate.SWAP(0, 1).update_quantum_state(sw)
    assert np.allclose(got, sw.get_vector(), atol=0)


def test_drop_qubit():
    n = 5
    s = _sv(n, 21)
    target, proj = [3, 0], [1, 0]
    d = state.drop_qubit(s, target, proj)
    t = s.get_vector().reshape((2,) * n)
    idx = [slice(None)] * n
    for qq, v in zip(target, proj):
        idx[n - 1 - qq] = v
    assert d.get_qubit_count() == 3
    assert np.allclose(d.get_vector(), t[tuple(idx)].reshape(-1), atol=0)


def test_partial_traces():
    n = 4
    s = _sv(n, 31)
    target = [2, 0]
    keep = [k for k in range(n) if k not in target]
    r = state.partial_trace(s, target)
    t = s.get_vector().reshape((2,) * n)
    kept_axes = [n - 1 - k for k in sorted(keep, reverse=True)]
    traced_axes = [n - 1 - k for k in target]
    m = np.transpose(t, kept_axes + traced_axes).reshape(1 << len(keep), -1)
    assert np.allclose(r.get_matrix(), m @ m.conj().T, atol=TOL)
    R = _dm(n, 33)
    rr = state.partial_trace(R, target)
    T = R.get_matrix().reshape((2,) * (2 * n))
    M = np.transpose(T, kept_axes + traced_axes + [n + a for a in kept_axes] + [n + a for a in traced_axes])
    M = M.reshape(1 << len(keep), 1 << len(target), 1 << len(keep), 1 << len(target))
    assert np.allclose(rr.get_matrix(), np.einsum("atbt->ab", M), atol=TOL)
    assert abs(np.trace(rr.get_matrix()) - 1.0) < TOL


def test_superposition_and_mixture():
    a, b = _sv(3, 41), _sv(3, 42)
    c1, c2 = 0.6 + 0.1j, -0.2j
    s = state.make_superposition(c1, a, c2, b)
    assert np.allclose(s.get_vector(), c1 * a.get_vector() + c2 * b.get_vector(), atol=TOL)
    R = _dm(3, 43)
    m = state.make_mixture(0.25, a, 0.75, R)
    want = 0.25 * np.outer(a.get_vector(), a.get_vector().conj()) + 0.75 * R.get_matrix()
    assert np.allclose(m.get_matrix(), want, atol=TOL)


def test_invalid_arguments():
    s = _sv(3, 1)
    with pytest.raises(RuntimeError):
        state.permutate_qubit(s, [0, 1])
    with pytest.raises(RuntimeError):
        state.permutate_qubit(s, [0, 0, 1])
    with pytest.raises(RuntimeError):
        state.drop_qubit(s, [0, 1, 2], [0, 0, 0])
    with pytest.raises(RuntimeError):
        state.drop_qubit(s, [0], [2])
    with pytest.raises(RuntimeError):
        state.partial_trace(s, [0, 1, 2])
