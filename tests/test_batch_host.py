"""Host-side validation of the batched-circuit entry points (no device work)."""
from __future__ import annotations

import numpy as np
import pytest

from qulacs_amd._lib import QulacsMI355Error
from qulacs_amd.batch import Batch, pauli_masks


def test_pauli_masks_match_reference_form():
    # X on 0, Y on 2, Z on 3 -> bit flip {0, 2}, phase flip {2, 3}, one Y
    assert pauli_masks([0, 2, 3], [1, 2, 3]) == (0b101, 0b1100, 1)


@pytest.mark.parametrize(
    "n,gates",
    [
        (0, []),
        (21, []),
        (3, [("dense", [3], np.eye(2))]),
        (3, [("dense", [1, 1], np.eye(4))]),
        (6, [("dense", [0, 1, 2, 3, 4], np.eye(32))]),
        (3, [("CZ", [0], None)]),
    ],
)
def test_add_circuit_rejects_bad_input(n, gates):
    b = Batch()
    with pytest.raises(QulacsMI355Error):
        b.add(n, gates, [])
    assert len(b) == 0


def test_term_beyond_qubits_rejected():
    b = Batch()
    with pytest.raises(QulacsMI355Error):
        b.add(3, [], [(1 << 3, 0, 0, 1.0)])


def test_compile_argument_checks_before_device_work():
    b = Batch()
    b.add(5, [("dense", [0, 1, 2, 3], np.eye(16))], [])
    with pytest.raises(QulacsMI355Error):
        b.compile(block_qubits=3)  # 4-qubit gate needs a 4-qubit block
    with pytest.raises(QulacsMI355Error):
        b.compile(block_qubits=5)


def test_fused_circuit_rejects_bad_arguments():
    from qulacs_amd.batch import FusedCircuit

    with pytest.raises(QulacsMI355Error):
        FusedCircuit(3, [])  # below 4 qubits
    with pytest.raises(QulacsMI355Error):
        FusedCircuit(10, [], local_qubits=7)
    with pytest.raises(QulacsMI355Error):
        FusedCircuit(10, [("dense", [0, 1, 2], np.eye(8))], block_qubits=2)
    fc = FusedCircuit(24, [("dense", [q], np.eye(2)) for q in range(24)] + [("CZ", [0, 23], None)])
    st = fc.stats()
    assert st["stages"] >= 2 and st["passes"] >= st["stages"]
    fc.close()


def test_fused_defaults_by_qubit_count():
    """Width and window 0 pick 3-qubit blocks and 12-qubit windows up to 19
    qubits, 4 and 13 above; a 4-qubit dense gate widens the default block."""
    from qulacs_amd.batch import FusedCircuit

    def stats(n, gates, **kw):
        fc = FusedCircuit(n, gates, **kw)
        try:
            return fc.stats()
        finally:
            fc.close()

    for n, (k, lq) in ((8, (3, 12)), (13, (3, 12)), (16, (3, 12)), (19, (3, 12)), (20, (4, 13)), (22, (4, 13))):
        gates = [("dense", [q], np.eye(2)) for q in range(n)] + [("CZ", [q, q + 1], None) for q in range(n - 1)]
        assert stats(n, gates) == stats(n, gates, block_qubits=k, local_qubits=lq), n
    wide = [("dense", [q], np.eye(2)) for q in range(16)] + [("dense", [1, 5, 9, 14], np.eye(16))]
    assert stats(16, wide) == stats(16, wide, block_qubits=4, local_qubits=12)


def test_server_rejects_bad_configuration():
    from qulacs_amd.serving import Server

    with pytest.raises(QulacsMI355Error):
        Server(hbm_budget_bytes=100)
    with pytest.raises(QulacsMI355Error):
        Server(hbm_budget_bytes=1 << 30, large_slots=0)
    with pytest.raises(QulacsMI355Error):
        Server(hbm_budget_bytes=1 << 30, small_max_qubits=17)


def test_adjoint_circuit_validation_is_host_only():
    from qulacs_amd.adjoint import AdjointCircuit, hardware_efficient_ansatz, heisenberg_terms

    gates, P = hardware_efficient_ansatz(20, 12)
    assert P == 480 and len(heisenberg_terms(20)) == 57
    c = AdjointCircuit(20, gates, P)
    c.close()
    with pytest.raises(QulacsMI355Error):
        AdjointCircuit(3, [("rot", [0], [4], 0, 1.0)], 1)  # Pauli id 4


def test_batch_window_setter_validation():
    from qulacs_amd import _lib

    b = Batch()
    with pytest.raises(QulacsMI355Error):
        _lib.call("qm_batch_set_local_qubits", b.h, 7)
    with pytest.raises(QulacsMI355Error):
        _lib.call("qm_batch_set_local_qubits", b.h, 14)
    _lib.call("qm_batch_set_local_qubits", b.h, 11)
    b.close()


def test_adjoint_pass_plan():
    """The adjoint groups gates into passes over 3-qubit register blocks (host
    planning, no device work): the 20-qubit, 12-layer hardware-efficient ansatz
    (480 rotations + 228 CZ) needs about one pass per three qubits per layer, a
    rotation with four X/Y qubits keeps one kernel per gate, and so does a state
    below three qubits."""
    pytest.importorskip("torch")
    from qulacs_amd.adjoint import AdjointCircuit, hardware_efficient_ansatz

    gates, P = hardware_efficient_ansatz(20, 12)
    st = AdjointCircuit(20, gates, P).stats()
    assert st["gates"] == 708 and 80 <= st["passes"] <= 84, st
    wide = AdjointCircuit(6, [("rot", [0, 1, 2, 3], [1, 2, 1, 1], 0, 1.0)], 1).stats()
    assert wide["passes"] == 0 and wide["gates"] == 1
    zrot = AdjointCircuit(6, [("rot", [0, 1, 2, 3], [3, 3, 3, 3], 0, 1.0), ("CZ", [0, 5], None)], 1).stats()
    assert zrot["passes"] == 1
    assert AdjointCircuit(2, [("rot", [0], [2], 0, 1.0)], 1).stats()["passes"] == 0
