"""Batched circuits on one device: many small circuits, each run from |0...0>
and reduced to the expectation value of its observable, in one kernel launch
(the qm_batch_* entry points of libqulacs_mi355).

    b = Batch()
    b.add(n, gates, terms)          # gates: (kind, qubits, matrix); terms: (bf, pf, rot, coef)
    b.compile(block_qubits=4)       # or b.autotune()
    energies = b.run()

Gate kinds are the strings "dense" (1..4 qubits, row-major 2^k x 2^k matrix with
bit i of the index <-> qubits[i]), "diag" (1..4 qubits, 2^k diagonal), "CZ" and
"CNOT" (control first).  Pauli terms use the reference's mask form
(src/csim/utility.cpp:12-43): bit-flip mask, phase-flip mask, number of Y's.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib

_KIND = {"dense": _lib.BATCH_DENSE, "diag": _lib.BATCH_DIAG, "CZ": _lib.BATCH_CZ, "CNOT": _lib.BATCH_CNOT}


def pauli_masks(targets, paulis):
    """(bit_flip, phase_flip, rot90) of a Pauli product (src/csim/utility.cpp:12-43)."""
    bf = pf = rot = 0
    for t, p in zip(targets, paulis):
        if p == 1:
            bf ^= 1 << t
        elif p == 2:
            bf ^= 1 << t
            pf ^= 1 << t
            rot += 1
        elif p == 3:
            pf ^= 1 << t
    return bf, pf, rot


class Batch:
    """A set of circuits compiled into one launch."""

    def __init__(self, device: int = 0):
        self.device = device
        self.h = _lib.call("qm_batch_create")
        self.sizes: list[int] = []
        self.block_qubits = 0
        self.keep_states = False

    def close(self) -> None:
        if self.h:
            _lib.call("qm_batch_destroy", self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __len__(self) -> int:
        return len(self.sizes)

    def add(self, n: int, gates, terms) -> int:
        """Add a circuit on n qubits; returns its index."""
        gates = list(gates)
        arr, keep = _marshal(gates)  # keep: the matrices stay alive until the call returns
        terms = list(terms)
        t = max(1, len(terms))
        bf = (C.c_ulonglong * t)(*[int(x[0]) for x in terms])
        pf = (C.c_ulonglong * t)(*[int(x[1]) for x in terms])
        rot = (C.c_uint * t)(*[int(x[2]) for x in terms])
        coef = (C.c_double * t)(*[float(x[3]) for x in terms])
        idx = _lib.call("qm_batch_add_circuit", self.h, int(n), arr, len(gates), bf, pf, rot, coef, len(terms))
        if idx < 0:
            raise _lib.QulacsMI355Error("qm_batch_add_circuit failed")
        self.sizes.append(int(n))
        return idx

    def compile(self, block_qubits: int = 4, keep_states: bool = False, local_qubits: int = 13) -> None:
        _lib.call("qm_batch_set_local_qubits", self.h, int(local_qubits))
        _lib.call("qm_batch_compile", self.h, int(block_qubits), 1 if keep_states else 0, self.device)
        self.block_qubits = block_qubits
        self.local_qubits = local_qubits
        self.keep_states = keep_states

    def run(self, stream=None) -> np.ndarray:
        out = np.zeros(max(1, len(self.sizes)), dtype=np.float64)
        _lib.call("qm_batch_run", self.h, out.ctypes.data_as(_lib.PD), stream, self.device)
        return out[: len(self.sizes)]

    def capture(self, stream) -> None:
        """Capture one run into a HIP graph on `stream` (from allocate_cuda_stream_host)."""
        _lib.call("qm_batch_capture", self.h, stream, self.device)

    def run_graph(self, stream) -> np.ndarray:
        out = np.zeros(max(1, len(self.sizes)), dtype=np.float64)
        _lib.call("qm_batch_run_graph", self.h, out.ctypes.data_as(_lib.PD), stream, self.device)
        return out[: len(self.sizes)]

    def run_async(self, stream=None) -> int:
        """Launch without waiting; returns the device address of the energies."""
        return _lib.call("qm_batch_run_async", self.h, stream, self.device)

    def state(self, index: int, stream=None) -> np.ndarray:
        out = np.zeros(1 << self.sizes[index], dtype=np.complex128)
        _lib.call("qm_batch_get_state", self.h, int(index), C.cast(out.ctypes.data, _lib.PC), stream, self.device)
        return out

    def stats(self) -> dict:
        v = [C.c_ulonglong() for _ in range(4)]
        _lib.call("qm_batch_stats", self.h, *[C.byref(x) for x in v])
        return dict(zip(("circuits", "stages", "passes", "subops"), (int(x.value) for x in v)))

    def autotune(self, candidates=(2, 3, 4), local_candidates=(12, 13), repeats: int = 2,
                 keep_states: bool = False) -> dict:
        """Compile with each (register-block size, LDS window) pair, time `repeats`
        runs of each and keep the fastest.  Returns {(block, window): best seconds}."""
        times = {}
        for k in candidates:
            for lq in local_candidates:
                try:
                    self.compile(k, keep_states, lq)
                except _lib.QulacsMI355Error:
                    continue  # a dense gate wider than k
                self.run()
                best = float("inf")
                for _ in range(repeats):
                    t0 = time.perf_counter()
                    self.run()
                    best = min(best, time.perf_counter() - t0)
                times[(k, lq)] = best
        if not times:
            raise _lib.QulacsMI355Error("autotune: no block size can hold the batch's gates")
        k, lq = min(times, key=times.get)
        self.compile(k, keep_states, lq)
        return times


class FusedCircuit:
    """One circuit applied in place to a device state by the stage scheduler
    (qm_circuit_*): one launch per stage of gates that fit a window of
    `local_qubits` qubits, register-blocked passes inside each stage."""

    def __init__(self, n: int, gates, block_qubits: int = 0, local_qubits: int = 0):
        """block_qubits / local_qubits 0: chosen by qubit count (qm_circuit_create)."""
        gates = list(gates)
        arr, self._keep = _marshal(gates)
        self.n = n
        self.h = _lib.call("qm_circuit_create", int(n), arr, len(gates), int(block_qubits), int(local_qubits))
        if not self.h:
            raise _lib.QulacsMI355Error("qm_circuit_create failed")

    def apply(self, state_ptr, stream=None, device: int = 0) -> None:
        _lib.call("qm_circuit_apply", self.h, state_ptr, 1 << self.n, stream, device)

    def stats(self) -> dict:
        v = [C.c_ulonglong() for _ in range(3)]
        _lib.call("qm_circuit_stats", self.h, *[C.byref(x) for x in v])
        return dict(zip(("stages", "passes", "subops"), (int(x.value) for x in v)))

    def close(self) -> None:
        if self.h:
            _lib.call("qm_circuit_destroy", self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def _marshal(gates):
    arr = (_lib.QmBatchGate * max(1, len(gates)))()
    keep = []
    for i, (kind, qubits, matrix) in enumerate(gates):
        if len(qubits) > 4 or kind not in _KIND:
            raise _lib.QulacsMI355Error(f"batch gate {i}: {kind} on {len(qubits)} qubits is not supported")
        g = arr[i]
        g.kind = _KIND[kind]
        g.qubit_count = len(qubits)
        for j, q in enumerate(qubits):
            g.qubits[j] = int(q)
        if matrix is not None:
            m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128)).ravel()
            keep.append(m)
            g.matrix = C.cast(m.ctypes.data, _lib.PC)
    return arr, keep
