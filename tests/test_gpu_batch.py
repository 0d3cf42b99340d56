"""Batched circuits (qm_batch_*) on the GPU against the oracle.

Bar: final states within 1e-10 relative L2 of the oracle; energies within
1e-10 * max(1, sum |c_k|); identical energies on repeated runs (the per-circuit
reduction has a fixed order)."""
from __future__ import annotations

import numpy as np
import pytest

from qulacs_amd.batch import Batch
from qulacs_amd.workloads import haar_unitary, many_small_circuits
from tests.batch_util import oracle_run, random_batch_circuit, random_terms
from tests.gpu_util import rel

pytestmark = pytest.mark.gpu


def _check(oracle, b, circuits, states=True):
    e = b.run()
    for i, (n, gates, terms) in enumerate(circuits):
        psi, want = oracle_run(oracle, n, gates, terms)
        scale = max(1.0, sum(abs(t[3]) for t in terms))
        if states:
            assert rel(b.state(i), psi) < 1e-10, (i, n, len(gates))
        assert abs(e[i] - want) < 1e-10 * scale, (i, n, e[i], want)
    return e


@pytest.mark.parametrize("block,local", [(4, 13), (3, 13), (2, 13), (4, 12), (3, 10)])
def test_random_circuits_all_sizes(oracle, block, local):
    rng = np.random.default_rng(100 + block + local)
    b = Batch()
    circuits = []
    for n in [1, 2, 3, 4, 5, 7, 9, 12, 13, 14, 15, 16, 17]:
        gates = random_batch_circuit(rng, n, 40 if n < 14 else 60, max_dense=block)
        terms = random_terms(rng, n, 6)
        circuits.append((n, gates, terms))
        b.add(n, gates, terms)
    b.compile(block_qubits=block, keep_states=True, local_qubits=local)
    print("block", block, "window", local, b.stats(), flush=True)
    e1 = _check(oracle, b, circuits)
    e2 = b.run()
    assert np.array_equal(e1, e2)
    b.close()


def test_empty_and_trivial_circuits(oracle):
    b = Batch()
    circuits = [(3, [], [(0, 1, 0, 0.5), (1, 0, 0, 2.0)]), (15, [], [(0, 1 << 14, 0, 1.0)]), (6, [], [])]
    for c in circuits:
        b.add(*c)
    b.compile(keep_states=True)
    e = _check(oracle, b, circuits)
    assert e[0] == 0.5 and e[1] == 1.0 and e[2] == 0.0


def test_many_small_circuits_workload_subset(oracle):
    cs = many_small_circuits(48, seed=77, n_range=(8, 16), layer_range=(2, 5))
    b = Batch()
    circuits = []
    for c in cs:
        b.add(c.n, c.gates, c.terms)
        circuits.append((c.n, c.gates, c.terms))
    times = b.autotune(repeats=1, keep_states=True)
    print("autotune", times, b.stats(), flush=True)
    _check(oracle, b, circuits)


def test_batch_errors():
    from qulacs_amd._lib import QulacsMI355Error

    b = Batch()
    b.add(4, [("dense", [0, 1, 2], np.eye(8))], [])
    with pytest.raises(QulacsMI355Error):
        b.compile(block_qubits=2)
    with pytest.raises(QulacsMI355Error):
        b.state(0)  # not compiled with keep_states
    b.compile(block_qubits=3)
    e = b.run()
    assert e.shape == (1,) and e[0] == 0.0


@pytest.mark.parametrize("n,block,local", [(4, 4, 13), (8, 2, 8), (13, 3, 13), (14, 4, 13), (16, 4, 10),
                                           (20, 3, 12), (22, 4, 13), (18, 3, 13), (16, 2, 13), (15, 2, 12)])
def test_fused_circuit_matches_oracle(oracle, n, block, local):
    from qulacs_amd import _lib
    from qulacs_amd.batch import FusedCircuit
    from tests.golden.gen import random_state
    from tests.gpu_util import DevState

    rng = np.random.default_rng(1000 + n)
    gates = random_batch_circuit(rng, n, 80, max_dense=block)
    psi = random_state(rng, n)
    want = psi.copy()
    for kind, q, m in gates:
        if kind == "dense":
            oracle.dense(q, m, want)
        elif kind == "diag":
            oracle.controlled_diag([], [], q, m, want)
        else:
            oracle.two(kind, q[0], q[1], want)
    fc = FusedCircuit(n, gates, block_qubits=block, local_qubits=local)
    print("fused", n, block, local, fc.stats(), flush=True)
    with DevState.from_host(psi) as s:
        fc.apply(s.ptr)
        got = s.download()
    assert rel(got, want) < 1e-10
    fc.close()


def test_graph_replays_match_eager_runs():
    from qulacs_amd import _lib

    rng = np.random.default_rng(55)
    b = Batch()
    for n in [4, 9, 13, 15]:
        b.add(n, random_batch_circuit(rng, n, 30), random_terms(rng, n, 5))
    b.compile(block_qubits=4)
    eager = b.run()
    stream = _lib.call("allocate_cuda_stream_host", 1, 0)
    try:
        b.capture(stream)
        for _ in range(4):
            assert np.array_equal(b.run_graph(stream), eager)
            assert np.array_equal(b.run(), eager)  # eager work between replays
    finally:
        _lib.call("release_cuda_stream_host", stream, 1, 0)
    b.close()


@pytest.mark.parametrize("block", [3, 4])
def test_fused_circuit_above_32_qubits_round_trip(block):
    """33 qubits (128 GiB): offsets of window qubits >= 32 exceed 32 bits.  U then
    U^dagger from |0...0> must return |0...0>: P(q = 0) = 1 for every probed qubit.
    Both register-block widths that prefetch the next tile (3: two blocks per thread,
    4: one 16-amplitude block)."""
    from qulacs_amd import _lib
    from qulacs_amd.batch import FusedCircuit

    n = 33
    rng = np.random.default_rng(3333)
    gates = [("dense", [32], haar_unitary(rng, 1)), ("CNOT", [32, 0], None), ("CZ", [31, 32], None),
             ("dense", [32, 5], haar_unitary(rng, 2))]
    gates += random_batch_circuit(rng, n, 20, max_dense=block)
    inverse = []
    for kind, q, m in reversed(gates):
        if kind == "dense":
            inverse.append((kind, q, np.conj(np.asarray(m)).T))
        elif kind == "diag":
            inverse.append((kind, q, np.conj(np.asarray(m))))
        else:
            inverse.append((kind, q, m))
    fc = FusedCircuit(n, gates + inverse, block_qubits=block)
    dim = 1 << n
    print("fused 33 qubits", fc.stats(), flush=True)
    try:
        st = _lib.call("allocate_quantum_state_host", dim, 0)
    except _lib.QulacsMI355Error as e:
        pytest.skip(f"no room for a 33-qubit state: {e}")
    try:
        _lib.call("initialize_quantum_state_host", st, dim, None, 0)
        fc.apply(st)
        for q in (0, 5, 17, 31, 32):
            p0 = _lib.call("M0_prob_host", q, st, dim, None, 0)
            assert abs(p0 - 1.0) < 1e-10, (q, p0)
        assert abs(_lib.call("state_norm_squared_host", st, dim, None, 0) - 1.0) < 1e-10
    finally:
        _lib.call("release_quantum_state_host", st, 0)
        fc.close()
