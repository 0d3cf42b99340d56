"""ctypes binding of libqulacs_mi355.so (the C ABI declared in include/qulacs_mi355.h).

The library is the only compute path: if it is missing or fails to load, this
module raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libqulacs_mi355.so"

ITYPE = C.c_ulonglong
UINT = C.c_uint
VP = C.c_void_p
D = C.c_double
PU = C.POINTER(C.c_uint)
PI = C.POINTER(C.c_ulonglong)
PD = C.POINTER(C.c_double)
PB = C.POINTER(C.c_ubyte)


class QmComplex(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


PC = C.POINTER(QmComplex)


class QmBatchGate(C.Structure):
    _fields_ = [("kind", C.c_uint), ("qubit_count", C.c_uint), ("qubits", C.c_uint * 4), ("matrix", PC)]


BATCH_DENSE, BATCH_DIAG, BATCH_CZ, BATCH_CNOT = 0, 1, 2, 3
ADJ_ROT = 4


class QmAdjointGate(C.Structure):
    _fields_ = [("kind", C.c_uint), ("qubit_count", C.c_uint), ("qubits", C.c_uint * 4), ("matrix", PC),
                ("paulis", C.c_uint * 4), ("param", C.c_uint), ("coef", C.c_double)]

# name -> (restype, argtypes)
_SIGS: dict[str, tuple] = {
    "qm_get_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "qm_clear_last_error": (None, []),
    "qm_set_synchronous": (None, [C.c_int]),
    "qm_get_synchronous": (C.c_int, []),
    "qm_synchronize": (None, [VP, UINT]),
    "qm_version": (C.c_char_p, []),
    "get_num_device": (C.c_int, []),
    "set_device": (None, [UINT]),
    "get_current_device": (C.c_int, []),
    "allocate_cuda_stream_host": (VP, [UINT, UINT]),
    "release_cuda_stream_host": (None, [VP, UINT, UINT]),
    "allocate_quantum_state_host": (VP, [ITYPE, UINT]),
    "initialize_quantum_state_host": (None, [VP, ITYPE, VP, UINT]),
    "release_quantum_state_host": (None, [VP, UINT]),
    "initialize_Haar_random_state_host": (None, [VP, ITYPE, VP, UINT]),
    "initialize_Haar_random_state_with_seed_host": (None, [VP, ITYPE, UINT, VP, UINT]),
    "set_computational_basis_host": (None, [ITYPE, VP, ITYPE, VP, UINT]),
    "copy_quantum_state_from_device_to_device": (None, [VP, VP, ITYPE, VP, UINT]),
    "copy_quantum_state_from_host_to_device": (None, [VP, VP, ITYPE, VP, UINT]),
    "copy_quantum_state_from_cppstate_host": (None, [VP, VP, ITYPE, VP, UINT]),
    "copy_quantum_state_from_device_to_host": (None, [VP, VP, ITYPE, VP, UINT]),
    "get_quantum_state_host": (None, [VP, VP, ITYPE, VP, UINT]),
    "print_quantum_state_host": (None, [VP, ITYPE, UINT]),
    # named gates
    **{f"{g}_gate_host": (None, [UINT, VP, ITYPE, VP, UINT])
       for g in ("H", "X", "Y", "Z", "S", "Sdag", "T", "Tdag", "sqrtX", "sqrtXdag", "sqrtY", "sqrtYdag", "P0", "P1")},
    "CZ_gate_host": (None, [UINT, UINT, VP, ITYPE, VP, UINT]),
    "CNOT_gate_host": (None, [UINT, UINT, VP, ITYPE, VP, UINT]),
    "SWAP_gate_host": (None, [UINT, UINT, VP, ITYPE, VP, UINT]),
    "RX_gate_host": (None, [UINT, D, VP, ITYPE, VP, UINT]),
    "RY_gate_host": (None, [UINT, D, VP, ITYPE, VP, UINT]),
    "RZ_gate_host": (None, [UINT, D, VP, ITYPE, VP, UINT]),
    "normalize_host": (None, [D, VP, ITYPE, VP, UINT]),
    "single_qubit_Pauli_gate_host": (None, [UINT, UINT, VP, ITYPE, VP, UINT]),
    "single_qubit_Pauli_rotation_gate_host": (None, [UINT, UINT, D, VP, ITYPE, VP, UINT]),
    "single_qubit_dense_matrix_gate_host": (None, [UINT, PC, VP, ITYPE, VP, UINT]),
    "single_qubit_diagonal_matrix_gate_host": (None, [UINT, PC, VP, ITYPE, VP, UINT]),
    "single_qubit_control_single_qubit_dense_matrix_gate_host": (None, [UINT, UINT, UINT, PC, VP, ITYPE, VP, UINT]),
    "single_qubit_phase_gate_host": (None, [UINT, QmComplex, VP, ITYPE, VP, UINT]),
    "multi_qubit_control_single_qubit_dense_matrix_gate_host": (None, [PU, PU, UINT, UINT, PC, VP, ITYPE, VP, UINT]),
    "penta_qubit_dense_matrix_gate_host": (None, [PU, PC, VP, ITYPE, VP, UINT]),
    "quad_qubit_dense_matrix_gate_host": (None, [PU, PC, VP, ITYPE, VP, UINT]),
    "triple_qubit_dense_matrix_gate_host": (None, [UINT, UINT, UINT, PC, VP, ITYPE, VP, UINT]),
    "double_qubit_dense_matrix_gate_host": (None, [UINT, UINT, PC, VP, ITYPE, VP, UINT]),
    "multi_qubit_Pauli_gate_Z_mask_host": (None, [ITYPE, VP, ITYPE, VP, UINT]),
    "multi_qubit_Pauli_gate_XZ_mask_host": (None, [ITYPE, ITYPE, UINT, UINT, VP, ITYPE, VP, UINT]),
    "multi_qubit_Pauli_rotation_gate_XZ_mask_host": (None, [ITYPE, ITYPE, UINT, UINT, D, VP, ITYPE, VP, UINT]),
    "multi_qubit_Pauli_rotation_gate_Z_mask_host": (None, [ITYPE, D, VP, ITYPE, VP, UINT]),
    "multi_qubit_Pauli_gate_partial_list_host": (None, [PU, PU, UINT, VP, ITYPE, VP, UINT]),
    "multi_qubit_Pauli_gate_whole_list_host": (None, [PU, UINT, VP, ITYPE, VP, UINT]),
    "multi_qubit_Pauli_rotation_gate_partial_list_host": (None, [PU, PU, UINT, D, VP, ITYPE, VP, UINT]),
    "multi_qubit_Pauli_rotation_gate_whole_list_host": (None, [PU, UINT, D, VP, ITYPE, VP, UINT]),
    "multi_qubit_dense_matrix_gate_host": (None, [PU, UINT, PC, VP, ITYPE, VP, UINT]),
    "single_qubit_control_multi_qubit_dense_matrix_gate_host": (None, [UINT, UINT, PU, UINT, PC, VP, ITYPE, VP, UINT]),
    "multi_qubit_control_multi_qubit_dense_matrix_gate_host": (None, [PU, PU, UINT, PU, UINT, PC, VP, ITYPE, VP, UINT]),
    "multi_qubit_diagonal_matrix_gate_host": (None, [PC, VP, ITYPE, VP, UINT]),
    "qm_multi_qubit_control_multi_qubit_diagonal_matrix_gate_host": (None, [PU, PU, UINT, PU, UINT, PC, VP, ITYPE, VP, UINT]),
    # reductions
    "state_norm_squared_cublas_host": (D, [VP, ITYPE]),
    "state_norm_squared_host": (D, [VP, ITYPE, VP, UINT]),
    "measurement_distribution_entropy_host": (D, [VP, ITYPE, VP, UINT]),
    "state_add_host": (None, [VP, VP, ITYPE, VP, UINT]),
    "state_multiply_host": (None, [QmComplex, VP, ITYPE, VP, UINT]),
    "inner_product_host": (QmComplex, [VP, VP, ITYPE, VP, UINT]),
    "expectation_value_single_qubit_Pauli_operator_host": (D, [UINT, UINT, VP, ITYPE, VP, UINT]),
    "M0_prob_host": (D, [UINT, VP, ITYPE, VP, UINT]),
    "M1_prob_host": (D, [UINT, VP, ITYPE, VP, UINT]),
    "marginal_prob_host": (D, [PU, PU, UINT, VP, ITYPE, VP, UINT]),
    "expectation_value_multi_qubit_Pauli_operator_XZ_mask_host": (D, [ITYPE, ITYPE, UINT, UINT, VP, ITYPE, VP, UINT]),
    "expectation_value_multi_qubit_Pauli_operator_Z_mask_host": (D, [ITYPE, VP, ITYPE, VP, UINT]),
    "expectation_value_multi_qubit_Pauli_operator_partial_list_host": (D, [PU, PU, UINT, VP, ITYPE, VP, UINT]),
    "expectation_value_multi_qubit_Pauli_operator_whole_list_host": (D, [PU, UINT, VP, ITYPE, VP, UINT]),
    "transition_amplitude_multi_qubit_Pauli_operator_XZ_mask_host": (QmComplex, [ITYPE, ITYPE, UINT, UINT, VP, VP, ITYPE, VP, UINT]),
    "transition_amplitude_multi_qubit_Pauli_operator_Z_mask_host": (QmComplex, [ITYPE, VP, VP, ITYPE, VP, UINT]),
    "transition_amplitude_multi_qubit_Pauli_operator_partial_list_host": (QmComplex, [PU, PU, UINT, VP, VP, ITYPE, VP, UINT]),
    "transition_amplitude_multi_qubit_Pauli_operator_whole_list_host": (QmComplex, [PU, UINT, VP, VP, ITYPE, VP, UINT]),
    "qm_state_add_with_coef_host": (None, [QmComplex, VP, VP, ITYPE, VP, UINT]),
    "qm_observable_expectation_host": (QmComplex, [PI, PI, PU, PC, UINT, VP, ITYPE, VP, UINT]),
    "qm_observable_expectation_noalloc_host": (QmComplex, [PI, PI, PU, PC, UINT, VP, ITYPE, VP, UINT]),
    "qm_sampling_host": (None, [PD, UINT, PI, VP, ITYPE, VP, UINT]),
    "qm_sampling_checked_host": (None, [PD, UINT, D, PI, PB, PU, VP, ITYPE, VP, UINT]),
    "qm_sampling_margin": (D, [UINT]),
    # density matrix
    "dm_initialize_quantum_state_host": (None, [VP, ITYPE, VP, UINT]),
    "dm_initialize_with_pure_state_host": (None, [VP, VP, ITYPE, VP, UINT]),
    "dm_single_qubit_dense_matrix_gate_host": (None, [UINT, PC, VP, ITYPE, VP, UINT]),
    "dm_multi_qubit_dense_matrix_gate_host": (None, [PU, UINT, PC, VP, ITYPE, VP, UINT]),
    "dm_multi_qubit_control_multi_qubit_dense_matrix_gate_host": (None, [PU, PU, UINT, PU, UINT, PC, VP, ITYPE, VP, UINT]),
    "dm_superoperator_host": (None, [PU, UINT, PC, VP, ITYPE, VP, UINT]),
    "dm_depolarizing_host": (None, [UINT, D, VP, ITYPE, VP, UINT]),
    "qm_dm_single_qubit_channels_host": (None, [PU, UINT, PC, VP, ITYPE, VP, UINT]),
    "dm_state_norm_squared_host": (D, [VP, ITYPE, VP, UINT]),
    "dm_measurement_distribution_entropy_host": (D, [VP, ITYPE, VP, UINT]),
    "dm_M0_prob_host": (D, [UINT, VP, ITYPE, VP, UINT]),
    "dm_marginal_prob_host": (D, [PU, PU, UINT, VP, ITYPE, VP, UINT]),
    "dm_expectation_value_multi_qubit_Pauli_operator_partial_list_host": (D, [PU, PU, UINT, VP, ITYPE, VP, UINT]),
    "dm_normalize_host": (None, [D, VP, ITYPE, VP, UINT]),
    "dm_sampling_host": (None, [PD, UINT, PI, VP, ITYPE, VP, UINT]),
    "qm_sampling_cdf_host": (None, [PD, VP, ITYPE, VP, UINT]),
    "dm_sampling_cdf_host": (None, [PD, VP, ITYPE, VP, UINT]),
    "dm_sampling_checked_host": (None, [PD, UINT, D, PI, PB, PU, VP, ITYPE, VP, UINT]),
    # batched circuits
    "qm_batch_create": (VP, []),
    "qm_batch_destroy": (None, [VP]),
    "qm_batch_add_circuit": (C.c_int, [VP, UINT, C.POINTER(QmBatchGate), UINT, PI, PI, PU, PD, UINT]),
    "qm_batch_compile": (C.c_int, [VP, UINT, C.c_int, UINT]),
    "qm_batch_run": (C.c_int, [VP, PD, VP, UINT]),
    "qm_batch_run_async": (VP, [VP, VP, UINT]),
    "qm_batch_get_state": (C.c_int, [VP, UINT, PC, VP, UINT]),
    "qm_batch_stats": (None, [VP, PI, PI, PI, PI]),
    "qm_batch_clear": (None, [VP]),
    "qm_batch_set_local_qubits": (C.c_int, [VP, UINT]),
    "qm_batch_capture": (C.c_int, [VP, VP, UINT]),
    "qm_batch_run_graph": (C.c_int, [VP, PD, VP, UINT]),
    "qm_batch_state_device_ptr": (VP, [VP, UINT]),
    "qm_circuit_create": (VP, [UINT, C.POINTER(QmBatchGate), UINT, UINT, UINT]),
    "qm_circuit_destroy": (None, [VP]),
    "qm_circuit_apply": (C.c_int, [VP, VP, ITYPE, VP, UINT]),
    "qm_circuit_stats": (None, [VP, PI, PI, PI]),
    "qm_adjoint_create": (VP, [UINT, C.POINTER(QmAdjointGate), UINT, UINT]),
    "qm_adjoint_destroy": (None, [VP]),
    "qm_adjoint_workspace_bytes": (ITYPE, [VP, UINT]),
    "qm_adjoint_stats": (None, [VP, PI, PI]),
    "qm_adjoint_run": (C.c_int, [VP, UINT, VP, PI, PI, PU, PD, UINT, VP, VP, VP, VP, UINT]),
    "qm_server_create": (VP, [ITYPE, UINT, UINT, UINT]),
    "qm_server_submit": (C.c_longlong, [VP, UINT, C.POINTER(QmBatchGate), UINT, PI, PI, PU, PD, UINT, PD, UINT]),
    "qm_server_wait": (C.c_int, [VP, C.c_longlong, PD, PI, PD]),
    "qm_server_wait_checked": (C.c_int, [VP, C.c_longlong, PD, PI, PB, PU, PD]),
    "qm_server_stats": (None, [VP, PI, PI, PI, PI, PI, PI]),
    "qm_server_destroy": (None, [VP]),
}

_lib = None


class QulacsMI355Error(RuntimeError):
    """Raised when an entry point records an error (invalid argument, HIP failure)."""


def lib() -> C.CDLL:
    """The loaded library, with argtypes/restype set for every entry point."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise ImportError(
                f"{_LIB_PATH} is missing: build it with `python -m qulacs_amd.build` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        l = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def symbol_names() -> list[str]:
    return list(_SIGS)


def check() -> None:
    """Raise QulacsMI355Error if the calling thread's last call recorded an error."""
    buf = C.create_string_buffer(512)
    code = lib().qm_get_last_error(buf, 512)
    if code != 0:
        lib().qm_clear_last_error()
        raise QulacsMI355Error(f"[{code}] {buf.value.decode(errors='replace')}")


def call(name: str, *args):
    """Call an entry point and raise on a recorded error."""
    r = getattr(lib(), name)(*args)
    check()
    return r
